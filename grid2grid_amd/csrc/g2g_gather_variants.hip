// One family of fused-gather variants: workgroup size G2G_THREADS and feature
// level G2G_FEAT (both set by the build, which compiles this file once per
// family so that the variants compile in parallel), every launch mode, bytes
// in flight and non-temporal hint.
#include "g2g_gather.hpp"

#if !defined(G2G_THREADS) || !defined(G2G_FEAT)
#error "G2G_THREADS and G2G_FEAT select the family"
#endif

namespace g2g {

#define G2G_FAMILY_DEF(TH, F)                                                                              \
    hipError_t launch_gather_##TH##_##F(const gather_args &a, const launch_cfg &c, int gl, hipStream_t s) { \
        return launch_family<TH, F>(a, c, gl, s);                                                          \
    }
#define G2G_FAMILY(TH, F) G2G_FAMILY_DEF(TH, F)
G2G_FAMILY(G2G_THREADS, G2G_FEAT)

}  // namespace g2g

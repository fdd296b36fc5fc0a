// The reference's scalapack2scalapack example (examples/scalapack2scalapack.cpp)
// and submatrix known-answer tests (examples/scalapack_submatrix.cpp,
// examples/parallel_scalapack_submatrix.cpp), written against the C++ facade
// (include/grid2grid) with P virtual ranks on one GPU instead of P MPI
// processes. Checks are bit-exact (the reference allows eps = 1e-6).
// Usage: scalapack2scalapack [m n ibm ibn fbm fbn pm pn]; no arguments runs
// the reference's CTest parameter sets (examples/CMakeLists.txt:20-30);
// --kat-only runs the host-only known-answer tests.
#include <grid2grid/scalapack_layout.hpp>
#include <grid2grid/transform.hpp>
#include <grid2grid/transformer.hpp>

#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace grid2grid;

#define HIPCHECK(x)                                                          \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            std::printf("HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            std::exit(2);                                                    \
        }                                                                    \
    } while (0)

// cantor_pairing(i, j) of the reference's test (cantor_mapping.hpp:14-17)
static double cantor(int i, int j) {
    const long s = i + j;
    return double(s * (s + 1) / 2 + j);
}

// Host image of one rank's block-cyclic local buffer, values f(i, j).
template <typename F>
static std::vector<double> local_values(scalapack::data_layout &l, int rank, F f) {
    std::vector<double> buf(scalapack::local_size(rank, l), -7.0);
    grid_layout<double> lay = get_scalapack_grid<double>(l, buf.data(), rank);
    for (int k = 0; k < lay.blocks.num_blocks(); ++k) {
        const block<double> &b = lay.blocks.get_block(k);
        for (int j = b.cols_interval.start; j < b.cols_interval.end; ++j)
            for (int i = b.rows_interval.start; i < b.rows_interval.end; ++i)
                b.data[(j - b.cols_interval.start) * b.stride + (i - b.rows_interval.start)] = f(i, j);
    }
    return buf;
}

struct rank_buffers {
    std::vector<double *> dev;
    std::vector<size_t> size;
    ~rank_buffers() {
        for (double *p : dev) (void)hipFree(p);
    }
};

// tuned: the layout pair is planned once (transform_plan), autotuned, and
// run again, then transform() runs a fresh plan of the same shape class
static bool run_case(int m, int n, int ibm, int ibn, int fbm, int fbn, int pm, int pn, char trans,
                     bool tuned = false) {
    const int P = pm * pn;
    scalapack::data_layout l1({m, n}, {ibm, ibn}, {pm, pn}, scalapack::column_major);
    const int fm = trans == 'N' ? m : n, fn = trans == 'N' ? n : m;
    scalapack::data_layout l2({fm, fn}, {fbm, fbn}, {pm, pn}, scalapack::column_major);
    rank_buffers b1, b2;
    std::vector<grid_layout<double>> per1, per2;
    for (int r = 0; r < P; ++r) {
        std::vector<double> h1 = local_values(l1, r, cantor);
        std::vector<double> h2(scalapack::local_size(r, l2), -1.0);
        double *d1 = nullptr, *d2 = nullptr;
        HIPCHECK(hipMalloc(&d1, sizeof(double) * (h1.size() + 1)));
        HIPCHECK(hipMalloc(&d2, sizeof(double) * (h2.size() + 1)));
        HIPCHECK(hipMemcpy(d1, h1.data(), sizeof(double) * h1.size(), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d2, h2.data(), sizeof(double) * h2.size(), hipMemcpyHostToDevice));
        b1.dev.push_back(d1);
        b1.size.push_back(h1.size());
        b2.dev.push_back(d2);
        b2.size.push_back(h2.size());
        per1.push_back(get_scalapack_grid<double>(l1, d1, r));
        per2.push_back(get_scalapack_grid<double>(l2, d2, r));
    }
    grid_layout<double> init = merge_ranks(per1), fin = merge_ranks(per2);
    bool ok = true;
    auto verify = [&]() {
        for (int r = 0; r < P; ++r) {
            std::vector<double> got(b2.size[r]);
            HIPCHECK(hipMemcpy(got.data(), b2.dev[r], sizeof(double) * got.size(), hipMemcpyDeviceToHost));
            std::vector<double> want = local_values(l2, r, [&](int i, int j) {
                return trans == 'N' ? cantor(i, j) : cantor(j, i);
            });
            for (size_t k = 0; k < got.size(); ++k) {
                // padding of the padded local buffers is never written: keeps -1
                const double w = want[k] == -7.0 ? -1.0 : want[k];
                if (got[k] != w) {
                    if (ok) std::printf("rank %d element %zu: %g instead of %g\n", r, k, got[k], w);
                    ok = false;
                }
            }
        }
    };
    auto reset = [&]() {  // the final buffers back to the sentinel values
        for (int r = 0; r < P; ++r)
            HIPCHECK(hipMemcpy(b2.dev[r], std::vector<double>(b2.size[r], -1.0).data(), sizeof(double) * b2.size[r],
                               hipMemcpyHostToDevice));
    };
    if (tuned) {
        transform_plan<double> plan(init, fin, trans);
        const g2g_tuning t = plan.autotune(nullptr, 1);
        std::printf("autotuned: %lld-byte work items, %d threads, column chunks %d\n", (long long)t.tile_bytes,
                    t.threads, t.column_chunks);
        reset();
        plan.execute();
        verify();
        // another configuration set by hand: 3-column chunks in destination
        // order, the last round of work items halved
        g2g_tuning u = plan.tuning();
        u.tile_bytes = 12288;
        u.column_chunks = 3;
        u.stream_order = G2G_ORDER_DESTINATION;
        u.row_chunks = 0;
        u.tail_split = 2;
        plan.set_tuning(u);
        const g2g_tuning v = plan.tuning();
        if (v.column_chunks != 3 || v.stream_order != G2G_ORDER_DESTINATION || v.tail_split != 2) {
            std::printf("set_tuning not reported back\n");
            ok = false;
        }
        reset();
        plan.execute();
        verify();
        reset();
    }
    int64_t hits0 = 0;
    detail::check(g2g_plan_cache_stats(&hits0, nullptr, nullptr));
    if (trans == 'N')
        transform<double>(init, fin);
    else
        transform<double>(init, fin, trans, 1.0, 0.0);
    verify();
    // the same layout pair again: served from the plan cache
    if (trans == 'N') {
        reset();
        transform<double>(init, fin);
        verify();
        int64_t hits1 = 0;
        detail::check(g2g_plan_cache_stats(&hits1, nullptr, nullptr));
        if (hits1 != hits0 + 1) {
            std::printf("plan cache: %lld hits after a repeated pair, %lld before\n", (long long)hits1,
                        (long long)hits0);
            ok = false;
        }
    }
    // the plan's launch replayed from its hipGraph, twice
    if (tuned) {
        hipStream_t st = nullptr;
        HIPCHECK(hipStreamCreate(&st));
        transform_plan<double> plan(init, fin, trans, st);
        virtual_comm comm;
        comm.stream = st;
        for (int rep = 0; rep < 2; ++rep) {
            reset();
            plan.execute_graph(1.0, 0.0, comm);
            verify();
        }
        HIPCHECK(hipStreamDestroy(st));
    }
    std::printf("m=%d n=%d init %dx%d final %dx%d grid %dx%d trans %c: %s\n", m, n, ibm, ibn, fbm, fbn, pm, pn,
                trans, ok ? "correct" : "NOT correct");
    return ok;
}

// known-answer tests of get_scalapack_grid (examples/scalapack_submatrix.cpp:26-78,
// examples/parallel_scalapack_submatrix.cpp:28-90), host pointers only
static bool submatrix_kats() {
    using namespace scalapack;
    bool ok = true;
    {
        std::vector<double> buffer(200 * 200);
        double *ptr = buffer.data();
        auto g = get_scalapack_grid<double>(200, {200, 200}, {1, 101}, {100, 100}, {32, 32}, {1, 1}, row_major,
                                            {0, 0}, ptr, 0);
        std::vector<int> rs{0, 32, 64, 96, 100}, cs{0, 28, 60, 92, 100};
        ok &= g.grid.grid().rows_split == rs && g.grid.grid().cols_split == cs;
        ok &= g.blocks.num_blocks() == 16;
        for (int k = 0; k < g.blocks.num_blocks() && ok; ++k) {
            const auto &b = g.blocks.get_block(k);
            const int i = k % 4, j = k / 4;
            ok &= b.coordinates.row == i && b.coordinates.col == j && b.stride == 200;
            ok &= b.data == ptr + rs[i] + 200 * (100 + cs[j]);
        }
    }
    {
        const int lld = 60;
        std::vector<double> buffer(lld * lld);
        double *ptr = buffer.data();
        const long want[4][4][3] = {{{0, 0, 20 + 7 * lld}, {2, 0, 27 + 7 * lld}, {0, 2, 20 + 13 * lld}, {2, 2, 27 + 13 * lld}},
                                    {{0, 1, 20}, {2, 1, 27}, {0, 3, 20 + 13 * lld}, {2, 3, 27 + 13 * lld}},
                                    {{1, 0, 7 * lld}, {3, 0, 27 + 7 * lld}, {1, 2, 13 * lld}, {3, 2, 27 + 13 * lld}},
                                    {{1, 1, 0}, {3, 1, 27}, {1, 3, 13 * lld}, {3, 3, 27 + 13 * lld}}};
        for (int rank = 0; rank < 4; ++rank) {
            auto g = get_scalapack_grid<double>(lld, {113, 56}, {21, 8}, {76, 37}, {27, 13}, {2, 2}, row_major,
                                                {0, 0}, ptr, rank);
            ok &= g.grid.grid().rows_split == std::vector<int>({0, 7, 34, 61, 76});
            ok &= g.grid.grid().cols_split == std::vector<int>({0, 6, 19, 32, 37});
            ok &= g.grid.owner(1, 2) == 2 && g.grid.owner(3, 3) == 3 && g.grid.owner(0, 1) == 1;
            ok &= g.blocks.num_blocks() == 4;
            for (int k = 0; k < 4 && ok; ++k) {
                const auto &b = g.blocks.get_block(k);
                ok &= b.coordinates.row == want[rank][k][0] && b.coordinates.col == want[rank][k][1];
                ok &= b.data == ptr + want[rank][k][2] && b.stride == lld;
            }
        }
    }
    std::printf("submatrix known-answer tests: %s\n", ok ? "SUCCESS" : "FAIL");
    return ok;
}

// The ScaLAPACK helpers (scalapack_layout.hpp:264-386 of the reference):
// global <-> local coordinates are inverse to each other and agree with the
// block placement of get_scalapack_grid; initialize_locally + validate round
// trip; rank_to_grid/rank_from_grid with a process-grid origin are inverse.
static bool helpers_case() {
    bool ok = true;
    const int m = 23, n = 17;
    for (scalapack::ordering ord : {scalapack::column_major, scalapack::row_major}) {
        scalapack::data_layout l({m, n}, {4, 3}, {2, 3}, ord);
        scalapack::matrix_grid mg(l.matrix_dimension, l.block_dimension);
        for (int gi = 0; gi < m; ++gi)
            for (int gj = 0; gj < n; ++gj) {
                scalapack::local_grid_coord lc = scalapack::local_coordinates(mg, l.rank_grid, {gi, gj});
                scalapack::elem_grid_coord g = scalapack::global_coordinates(mg, l.rank_grid, lc);
                ok &= g.row == gi && g.col == gj;
                // the owner's get_scalapack_grid block holds (gi, gj) at the same buffer offset
                const int rank = scalapack::rank_from_grid(lc.rank_coord, l.rank_grid, ord);
                std::vector<double> buf(scalapack::local_size(rank, l));
                grid_layout<double> lay = get_scalapack_grid<double>(l, buf.data(), rank);
                scalapack::local_blocks lb = scalapack::get_local_blocks(mg, l.rank_grid, lc.rank_coord);
                bool found = false;
                for (int k = 0; k < lay.blocks.num_blocks(); ++k) {
                    const block<double> &b = lay.blocks.get_block(k);
                    if (gi < b.rows_interval.start || gi >= b.rows_interval.end || gj < b.cols_interval.start ||
                        gj >= b.cols_interval.end)
                        continue;
                    const double *p = b.data + (gj - b.cols_interval.start) * b.stride + (gi - b.rows_interval.start);
                    found = p == buf.data() + (size_t)lc.el_coord.col * lb.stride + lc.el_coord.row;
                }
                ok &= found;
            }
        for (int r = 0; r < 6; ++r) {
            std::vector<double> buf(scalapack::local_size(r, l));
            scalapack::initialize_locally(buf.data(), cantor, r, l);
            ok &= scalapack::validate(cantor, buf, r, l);
            scalapack::rank_grid_coord src{1, 2};
            scalapack::rank_grid_coord c = scalapack::rank_to_grid(r, l.rank_grid, ord, src);
            ok &= scalapack::rank_from_grid(c, l.rank_grid, ord, src) == r;
        }
    }
    // the reference's row test is gi > m: the padding row gi == m of a
    // partial last block maps to (m, gj), one row further is (-1, -1)
    scalapack::matrix_grid mg({10, 4}, {4, 4});
    scalapack::elem_grid_coord a = scalapack::global_coordinates(mg, {1, 1}, {{10, 1}, {0, 0}});
    scalapack::elem_grid_coord b = scalapack::global_coordinates(mg, {1, 1}, {{11, 1}, {0, 0}});
    ok &= a.row == 10 && a.col == 1 && b.row == -1 && b.col == -1;
    std::printf("scalapack helpers: %s\n", ok ? "SUCCESS" : "FAIL");
    return ok;
}

// transformer<T>: several pairs, mixed 'N'/'T', one launch
static bool transformer_case() {
    const int P = 4;
    scalapack::data_layout a1({30, 20}, {4, 3}, {2, 2}, scalapack::column_major);
    scalapack::data_layout a2({30, 20}, {7, 5}, {2, 2}, scalapack::column_major);
    scalapack::data_layout c1({18, 25}, {5, 5}, {2, 2}, scalapack::column_major);
    scalapack::data_layout c2({25, 18}, {4, 6}, {2, 2}, scalapack::column_major);
    rank_buffers bufs;
    auto make = [&](scalapack::data_layout &l, bool fill, std::vector<double *> &ptrs) {
        std::vector<grid_layout<double>> per;
        for (int r = 0; r < P; ++r) {
            std::vector<double> h = fill ? local_values(l, r, cantor) : std::vector<double>(scalapack::local_size(r, l), -1.0);
            double *d = nullptr;
            HIPCHECK(hipMalloc(&d, sizeof(double) * (h.size() + 1)));
            HIPCHECK(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
            bufs.dev.push_back(d);
            ptrs.push_back(d);
            per.push_back(get_scalapack_grid<double>(l, d, r));
        }
        return merge_ranks(per);
    };
    std::vector<double *> pa1, pa2, pc1, pc2;
    grid_layout<double> A1 = make(a1, true, pa1), A2 = make(a2, false, pa2);
    grid_layout<double> C1 = make(c1, true, pc1), C2 = make(c2, false, pc2);
    transformer<double> t;
    t.schedule(A1, A2);
    t.schedule(C1, C2, 'T', 1.0, 0.0);
    t.transform();
    bool ok = true;
    for (int r = 0; r < P; ++r) {
        std::vector<double> g2(scalapack::local_size(r, a2)), g4(scalapack::local_size(r, c2));
        HIPCHECK(hipMemcpy(g2.data(), pa2[r], sizeof(double) * g2.size(), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(g4.data(), pc2[r], sizeof(double) * g4.size(), hipMemcpyDeviceToHost));
        auto w2 = local_values(a2, r, cantor);
        auto w4 = local_values(c2, r, [](int i, int j) { return cantor(j, i); });
        for (size_t k = 0; k < g2.size(); ++k) ok &= g2[k] == (w2[k] == -7.0 ? -1.0 : w2[k]);
        for (size_t k = 0; k < g4.size(); ++k) ok &= g4[k] == (w4[k] == -7.0 ? -1.0 : w4[k]);
    }
    std::printf("transformer (N + T pairs in one launch): %s\n", ok ? "correct" : "NOT correct");
    return ok;
}

// grid_layout::scale_by and block::scale_by / local_element on device
// memory (grid_layout.hpp:48-56, block.cpp:228-306): every element doubled
// (exact), padding untouched, one element read and written through the proxy.
static bool scale_case() {
    const int P = 4;
    scalapack::data_layout l({30, 20}, {4, 3}, {2, 2}, scalapack::column_major);
    rank_buffers bufs;
    std::vector<grid_layout<double>> per;
    std::vector<std::vector<double>> host;
    for (int r = 0; r < P; ++r) {
        host.push_back(local_values(l, r, cantor));
        double *d = nullptr;
        HIPCHECK(hipMalloc(&d, sizeof(double) * (host[r].size() + 1)));
        HIPCHECK(hipMemcpy(d, host[r].data(), sizeof(double) * host[r].size(), hipMemcpyHostToDevice));
        bufs.dev.push_back(d);
        bufs.size.push_back(host[r].size());
        per.push_back(get_scalapack_grid<double>(l, d, r));
    }
    grid_layout<double> lay = merge_ranks(per);
    lay.scale_by(2.0);
    lay.scale_by(1.0);  // no-op
    bool ok = true;
    for (int r = 0; r < P; ++r) {
        std::vector<double> got(bufs.size[r]);
        HIPCHECK(hipMemcpy(got.data(), bufs.dev[r], sizeof(double) * got.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < got.size(); ++k)
            ok &= got[k] == (host[r][k] == -7.0 ? -7.0 : 2.0 * host[r][k]);
    }
    block<double> &b = lay.blocks.get_block(3);
    const int gi = b.rows_interval.start + 1, gj = b.cols_interval.start + 2;
    ok &= double(b.local_element(1, 2)) == 2.0 * cantor(gi, gj);
    b.local_element(1, 2) = 0.5;
    ok &= double(b.local_element(1, 2)) == 0.5;
    b.scale_by(4.0);
    ok &= double(b.local_element(1, 2)) == 2.0 && double(b.local_element(0, 0)) ==
          8.0 * cantor(b.rows_interval.start, b.cols_interval.start);
    bool refused = false;
    block<double> t = b;
    t.transpose_or_conjugate('T');
    try {
        t.scale_by(3.0);
    } catch (const std::runtime_error &) {
        refused = true;
    }
    ok &= refused;
    std::printf("scale_by / local_element: %s\n", ok ? "correct" : "NOT correct");
    return ok;
}

// communication_volume against a count over elements, and optimal_reordering:
// relabelling the final grid's ranks with the returned permutation lowers the
// remote volume by exactly the adjusted weights of the matched pairs
// (ranks_reordering.cpp:17-34).
static bool volume_case() {
    scalapack::data_layout a({30, 20}, {4, 3}, {2, 2}, scalapack::column_major);
    scalapack::data_layout b({30, 20}, {7, 5}, {2, 2}, scalapack::row_major);
    std::vector<double> dummy(scalapack::local_size(0, a) + scalapack::local_size(0, b));
    assigned_grid2D gi = get_scalapack_grid<double>(a, dummy.data(), 0).grid;
    assigned_grid2D gf = get_scalapack_grid<double>(b, dummy.data(), 0).grid;
    auto owner_at = [](const assigned_grid2D &g, int i, int j) {
        int bi = 0, bj = 0;
        while (g.grid().rows_split[bi + 1] <= i) ++bi;
        while (g.grid().cols_split[bj + 1] <= j) ++bj;
        return g.owner(bi, bj);
    };
    comm_volume v0 = communication_volume(gi, gf);
    comm_volume brute;
    for (int i = 0; i < 30; ++i)
        for (int j = 0; j < 20; ++j) brute.volume[edge_t{owner_at(gi, i, j), owner_at(gf, i, j)}.sorted()] += 1;
    bool ok = v0.volume == brute.volume;
    bool reordered = false;
    std::vector<int> perm = optimal_reordering(v0, 4, reordered);
    long gain = 0;
    for (int r = 0; r < 4; ++r) {
        const int s = perm[r];
        ok &= perm[s] == r;
        if (r < s) gain += (long)v0.volume[edge_t{r, s}] - v0.volume[edge_t{r, r}] - v0.volume[edge_t{s, s}];
    }
    gf.reorder_ranks(perm);
    comm_volume v1 = communication_volume(gi, gf);
    ok &= reordered && gain > 0 && (long)v1.total_volume() == (long)v0.total_volume() - gain;
    std::printf("communication_volume + optimal_reordering (remote %zu -> %zu): %s\n", v0.total_volume(),
                v1.total_volume(), ok ? "correct" : "NOT correct");
    return ok;
}

int main(int argc, char **argv) {
    bool ok = submatrix_kats();
    ok &= helpers_case();
    if (argc == 2 && std::string(argv[1]) == "--kat-only") return !ok;  // host-only part
    if (argc == 9) {
        int v[8];
        for (int k = 0; k < 8; ++k) v[k] = std::atoi(argv[k + 1]);
        ok &= run_case(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], 'N');
        return !ok;
    }
    const int cases[4][8] = {{10, 10, 2, 2, 5, 5, 2, 2}, {10, 10, 2, 3, 3, 5, 2, 2},
                             {10, 12, 2, 3, 2, 6, 2, 2}, {10, 12, 2, 3, 3, 5, 2, 2}};
    for (const auto &c : cases) ok &= run_case(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], 'N');
    ok &= run_case(64, 48, 8, 8, 5, 7, 2, 2, 'T');
    ok &= run_case(700, 500, 24, 40, 56, 32, 2, 2, 'N', true);
    ok &= transformer_case();
    ok &= volume_case();
    ok &= scale_case();
    std::printf("%s\n", ok ? "ALL CORRECT" : "FAILURES");
    return !ok;
}

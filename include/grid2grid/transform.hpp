// grid2grid::transform -- the reference's entry points
// (src/grid2grid/transform.hpp:13-43) on one MI355X through the C ABI of
// libgrid2grid_amd.so. MPI_Comm becomes virtual_comm: the P ranks of the
// layouts are virtual ranks whose blocks all live in this GPU's memory.
//
// As in the reference, transform() returns when the final layout holds the
// data (unless virtual_comm::synchronize is false) and the overloads with
// `trans` leave the initial layout transposed/conjugated (transform.cpp:151,
// 197). Unlike the reference, which plans every call, the fused single-pair
// overloads plan a layout pair once: g2g_transform_layouts keeps the plan
// (grid cover, pieces, offsets) in a process-wide cache keyed by the layouts'
// descriptors. transform_plan holds the plan of a layout pair explicitly.
#pragma once

#include <grid2grid_amd.h>

#include <grid2grid/comm_volume.hpp>
#include <grid2grid/grid_layout.hpp>
#include <grid2grid/ranks_reordering.hpp>

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace grid2grid {

enum class exchange_mode {
    fused,   // one gather kernel: pack, exchange and unpack fused
    staged   // pack -> exchange -> unpack through per-rank buffers, as the reference does
};

struct virtual_comm {
    void *stream = nullptr;  // hipStream_t of the work
    exchange_mode mode = exchange_mode::fused;
    bool synchronize = true;  // wait for the data before returning
};

namespace detail {

// g2g_layout view of a grid_layout (storage orientation required).
template <typename T>
struct layout_desc {
    std::vector<int32_t> rs, cs, owners;
    std::vector<g2g_block> blocks;
    g2g_layout l{};

    // Built on every call of the drop-in transform(): one pass over the
    // owners and one over the blocks, the splits read directly (config-2
    // sized layouts, 74 000 blocks, take well under a millisecond).
    explicit layout_desc(const grid_layout<T> &lay) {
        const grid2D &g = lay.grid.grid();
        rs.assign(g.rows_split.begin(), g.rows_split.end());
        cs.assign(g.cols_split.begin(), g.cols_split.end());
        owners.resize(size_t(g.n_rows) * g.n_cols);
        int32_t *o = owners.data();
        for (int i = 0; i < g.n_rows; ++i)
            for (int j = 0; j < g.n_cols; ++j) *o++ = lay.grid.owner(i, j);
        const int n = lay.blocks.num_blocks();
        blocks.reserve(size_t(n));
        for (int k = 0; k < n; ++k) {
            const block<T> &b = lay.blocks.get_block(k);
            if (b.transpose_on_copy)
                throw std::runtime_error("grid2grid: layout is already transposed; pass trans instead");
            const int r = b.coordinates.row, c = b.coordinates.col;
            if (r < 0 || r >= g.n_rows || c < 0 || c >= g.n_cols ||
                b.rows_interval != interval(rs[size_t(r)], rs[size_t(r) + 1]) ||
                b.cols_interval != interval(cs[size_t(c)], cs[size_t(c) + 1]))
                throw std::runtime_error("grid2grid: a block does not match its grid cell");
            blocks.push_back(g2g_block{r, c, b.stride, (void *)b.data});
        }
        l.n_rows_split = int32_t(rs.size());
        l.n_cols_split = int32_t(cs.size());
        l.rows_split = rs.data();
        l.cols_split = cs.data();
        l.owners = owners.data();
        l.n_ranks = lay.num_ranks();
        l.n_blocks = int32_t(blocks.size());
        l.blocks = blocks.data();
    }
};

inline char upper(char c) { return c >= 'a' && c <= 'z' ? char(c - 'a' + 'A') : c; }

}  // namespace detail

// A planned transform of one layout pair, reusable across calls.
template <typename T>
class transform_plan {
  public:
    transform_plan(const grid_layout<T> &initial, const grid_layout<T> &final_, char trans = 'N',
                   void *stream = nullptr) {
        detail::layout_desc<T> di(initial), df(final_);
        detail::check(g2g_plan_create(&di.l, &df.l, detail::dtype_of<T>::value, detail::upper(trans), stream, &p_));
    }
    transform_plan(const transform_plan &) = delete;
    transform_plan &operator=(const transform_plan &) = delete;
    ~transform_plan() {
        if (p_) g2g_plan_destroy(p_);
    }
    // final = beta*final + alpha*op(initial); (1, 0) is a plain copy
    void execute(T alpha = T{1}, T beta = T{0}, const virtual_comm &comm = {}) {
        const bool plain = alpha == T{1} && beta == T{0};
        const void *a = plain ? nullptr : &alpha, *b = plain ? nullptr : &beta;
        detail::check(comm.mode == exchange_mode::fused ? g2g_transform(p_, a, b, comm.stream)
                                                        : g2g_transform_staged(p_, a, b, comm.stream));
        if (comm.synchronize) detail::check(g2g_synchronize(comm.stream));
    }
    // the same launch replayed from the plan's hipGraph (g2g_transform_graph);
    // comm.stream must not be the null stream
    void execute_graph(T alpha = T{1}, T beta = T{0}, const virtual_comm &comm = {}) {
        const bool plain = alpha == T{1} && beta == T{0};
        const void *a = plain ? nullptr : &alpha, *b = plain ? nullptr : &beta;
        detail::check(g2g_transform_graph(p_, a, b, comm.stream));
        if (comm.synchronize) detail::check(g2g_synchronize(comm.stream));
    }
    g2g_plan_info info() const {
        g2g_plan_info i;
        detail::check(g2g_plan_get_info(p_, &i));
        return i;
    }
    // Times the compiled variants on this layout pair (the final layout is
    // written meanwhile) and keeps the fastest; plans created later for the
    // same shape class start from it (g2g_plan_autotune).
    g2g_tuning autotune(void *stream = nullptr, int reps = 3) {
        g2g_tuning t;
        detail::check(g2g_plan_autotune(p_, stream, reps, &t));
        return t;
    }
    // The launch configuration (g2g_plan_get_tuning / g2g_plan_set_tuning):
    // work-item size, kernel variant, stream (column chunks, row chunks,
    // order) and tail split.
    g2g_tuning tuning() const {
        g2g_tuning t;
        detail::check(g2g_plan_get_tuning(p_, &t));
        return t;
    }
    void set_tuning(const g2g_tuning &t) { detail::check(g2g_plan_set_tuning(p_, &t)); }
    g2g_plan handle() const { return p_; }

  private:
    g2g_plan p_ = nullptr;
};

namespace detail {

// One pair through the plan cache (fused form) or a plan of its own (staged form).
template <typename T>
void transform_pair(grid_layout<T> &initial_layout, grid_layout<T> &final_layout, char trans, const T &alpha,
                    const T &beta, const virtual_comm &comm) {
    trans = upper(trans);
    const bool plain = alpha == T{1} && beta == T{0};
    if (comm.mode == exchange_mode::staged) {
        transform_plan<T> p(initial_layout, final_layout, trans, comm.stream);
        p.execute(alpha, beta, comm);
        return;
    }
    layout_desc<T> di(initial_layout), df(final_layout);
    check(g2g_transform_layouts(&di.l, &df.l, dtype_of<T>::value, trans, plain ? nullptr : &alpha,
                                plain ? nullptr : &beta, comm.stream));
    if (comm.synchronize) check(g2g_synchronize(comm.stream));
}

}  // namespace detail

// final_layout = initial_layout
template <typename T>
void transform(grid_layout<T> &initial_layout, grid_layout<T> &final_layout, const virtual_comm &comm = {}) {
    detail::transform_pair<T>(initial_layout, final_layout, 'N', T{1}, T{0}, comm);
}

// final_layout = beta * final_layout + alpha * op(initial_layout); leaves
// initial_layout transposed/conjugated like the reference.
template <typename T>
void transform(grid_layout<T> &initial_layout, grid_layout<T> &final_layout, const char trans, const T alpha,
               const T beta, const virtual_comm &comm = {}) {
    detail::transform_pair<T>(initial_layout, final_layout, trans, alpha, beta, comm);
    initial_layout.transpose_or_conjugate(trans);
}

// Several layout pairs in one persistent launch.
template <typename T>
void transform(std::vector<layout_ref<T>> &initial_layouts, std::vector<layout_ref<T>> &final_layouts,
               const char *trans, const T *alpha, const T *beta, const virtual_comm &comm = {}) {
    if (initial_layouts.size() != final_layouts.size())
        throw std::runtime_error("grid2grid: initial and final layout lists differ in length");
    const size_t n = initial_layouts.size();
    std::vector<g2g_plan> plans;
    std::vector<const void *> al(n), be(n);
    int rc = G2G_OK;
    for (size_t k = 0; k < n && rc == G2G_OK; ++k) {
        detail::layout_desc<T> di(initial_layouts[k].get()), df(final_layouts[k].get());
        g2g_plan p = nullptr;
        rc = g2g_plan_create(&di.l, &df.l, detail::dtype_of<T>::value, trans ? detail::upper(trans[k]) : 'N',
                             comm.stream, &p);
        if (rc == G2G_OK) plans.push_back(p);
        al[k] = alpha ? (const void *)&alpha[k] : nullptr;
        be[k] = beta ? (const void *)&beta[k] : nullptr;
    }
    g2g_batch b = nullptr;
    if (rc == G2G_OK)
        rc = g2g_batch_create(int(n), plans.data(), alpha ? al.data() : nullptr, beta ? be.data() : nullptr,
                              comm.stream, &b);
    if (rc == G2G_OK) rc = g2g_batch_execute(b, comm.stream);
    if (rc == G2G_OK && comm.synchronize) rc = g2g_synchronize(comm.stream);
    std::string err = rc == G2G_OK ? "" : g2g_last_error();
    if (b) g2g_batch_destroy(b);
    for (g2g_plan p : plans) g2g_plan_destroy(p);
    if (rc != G2G_OK) throw std::runtime_error("grid2grid: " + err);
    if (trans)
        for (size_t k = 0; k < n; ++k) initial_layouts[k].get().transpose_or_conjugate(trans[k]);
}

template <typename T>
void transform(std::vector<layout_ref<T>> &initial_layouts, std::vector<layout_ref<T>> &final_layouts,
               const virtual_comm &comm = {}) {
    transform<T>(initial_layouts, final_layouts, nullptr, nullptr, nullptr, comm);
}

// Elements exchanged between each pair of ranks when data moves from g_init
// to g_final (transform.hpp:47, transform.cpp:9-39), computed on the device
// (g2g_comm_volume). Owners are taken after any rank reordering, as
// assigned_grid2D::owner gives them.
inline comm_volume communication_volume(assigned_grid2D &g_init, assigned_grid2D &g_final,
                                        void *stream = nullptr) {
    struct grid_desc {
        std::vector<int32_t> rs, cs, owners;
        g2g_layout l{};
        explicit grid_desc(const assigned_grid2D &g) {
            const grid2D &gr = g.grid();
            rs.assign(gr.rows_split.begin(), gr.rows_split.end());
            cs.assign(gr.cols_split.begin(), gr.cols_split.end());
            owners.resize(size_t(gr.n_rows) * gr.n_cols);
            for (int i = 0; i < gr.n_rows; ++i)
                for (int j = 0; j < gr.n_cols; ++j) owners[size_t(i) * gr.n_cols + j] = g.owner(i, j);
            l.n_rows_split = int32_t(rs.size());
            l.n_cols_split = int32_t(cs.size());
            l.rows_split = rs.data();
            l.cols_split = cs.data();
            l.owners = owners.data();
            l.n_ranks = g.num_ranks();
        }
    };
    grid_desc di(g_init), df(g_final);
    int64_t n = 0;
    detail::check(g2g_comm_volume(&di.l, &df.l, nullptr, 0, &n, stream));
    std::vector<g2g_edge> edges((size_t)n);
    detail::check(g2g_comm_volume(&di.l, &df.l, edges.data(), n, &n, stream));
    comm_volume::volume_t v;
    for (const g2g_edge &e : edges) {
        if (e.weight > 0x7fffffffll)
            throw std::runtime_error("grid2grid: communication volume of a rank pair exceeds int");
        v[edge_t{e.src, e.dest}] = int(e.weight);
    }
    return comm_volume(std::move(v));
}

}  // namespace grid2grid

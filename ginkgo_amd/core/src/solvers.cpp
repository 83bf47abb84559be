// Jacobi, Cg, Bicgstab and Ir on the Mi355xExecutor.
//
// Cg has three paths.  The fused path (one right-hand side, classical CSR,
// identity or scalar Jacobi preconditioner) runs two kernels per iteration
// and evaluates the stopping criteria on the device; the host enqueues
// iterations ahead in chunks and polls the device flags through pinned
// memory one chunk behind, so it never stalls the stream inside the loop.
// The fused block path does the same for several right-hand sides over any
// Csr.  The generic path follows core/solver/cg.cpp step by step with the
// reference-semantics kernels and handles any LinOp system matrix.
// Bicgstab has the fused path for one right-hand side (five kernels per
// iteration, both checks on the device) and the generic path otherwise.
#include <cmath>

#include "ginkgo_amd/dispatch.hpp"
#include "ginkgo_amd/solver.hpp"

namespace gko {

namespace {

template <typename T>
struct sk;

template <>
struct sk<double> {
    static constexpr auto jacobi_generate = gkomi_jacobi_generate_f64;
    static constexpr auto jacobi_apply = gkomi_jacobi_apply_f64;
    static constexpr auto jacobi_adv = gkomi_jacobi_advanced_apply_f64;
    static constexpr auto cg_init = gkomi_cg_initialize_f64;
    static constexpr auto cg_step_1 = gkomi_cg_step_1_f64;
    static constexpr auto cg_step_2 = gkomi_cg_step_2_f64;
    static constexpr auto cg_fused_start = gkomi_cg_fused_start_f64;
    static constexpr auto cg_fused_iterate = gkomi_cg_fused_iterate_f64;
    static constexpr auto block_cg_start = gkomi_block_cg_start_f64;
    static constexpr auto block_cg_iterate = gkomi_block_cg_iterate_f64;
    static constexpr auto bicg_init = gkomi_bicgstab_initialize_f64;
    static constexpr auto bicg_step_1 = gkomi_bicgstab_step_1_f64;
    static constexpr auto bicg_step_2 = gkomi_bicgstab_step_2_f64;
    static constexpr auto bicg_step_3 = gkomi_bicgstab_step_3_f64;
    static constexpr auto bicg_finalize = gkomi_bicgstab_finalize_f64;
    static constexpr auto bicg_fused_start = gkomi_bicgstab_fused_start_f64;
    static constexpr auto bicg_fused_iterate = gkomi_bicgstab_fused_iterate_f64;
    static constexpr auto check = gkomi_criteria_check_f64;
    static constexpr auto fill = gkomi_fill_f64;
};

template <>
struct sk<float> {
    static constexpr auto jacobi_generate = gkomi_jacobi_generate_f32;
    static constexpr auto jacobi_apply = gkomi_jacobi_apply_f32;
    static constexpr auto jacobi_adv = gkomi_jacobi_advanced_apply_f32;
    static constexpr auto cg_init = gkomi_cg_initialize_f32;
    static constexpr auto cg_step_1 = gkomi_cg_step_1_f32;
    static constexpr auto cg_step_2 = gkomi_cg_step_2_f32;
    static constexpr auto cg_fused_start = gkomi_cg_fused_start_f32;
    static constexpr auto cg_fused_iterate = gkomi_cg_fused_iterate_f32;
    static constexpr auto block_cg_start = gkomi_block_cg_start_f32;
    static constexpr auto block_cg_iterate = gkomi_block_cg_iterate_f32;
    static constexpr auto bicg_init = gkomi_bicgstab_initialize_f32;
    static constexpr auto bicg_step_1 = gkomi_bicgstab_step_1_f32;
    static constexpr auto bicg_step_2 = gkomi_bicgstab_step_2_f32;
    static constexpr auto bicg_step_3 = gkomi_bicgstab_step_3_f32;
    static constexpr auto bicg_finalize = gkomi_bicgstab_finalize_f32;
    static constexpr auto bicg_fused_start = gkomi_bicgstab_fused_start_f32;
    static constexpr auto bicg_fused_iterate = gkomi_bicgstab_fused_iterate_f32;
    static constexpr auto check = gkomi_criteria_check_f32;
    static constexpr auto fill = gkomi_fill_f32;
};

template <typename V>
using Vec = matrix::Dense<V>;

template <typename V>
std::unique_ptr<Vec<V>> scalar(std::shared_ptr<const Executor> exec, V value, size_type cols = 1)
{
    auto s = Vec<V>::create(exec, dim<2>{1, cols});
    s->fill(value);
    return s;
}

// Pinned host slots for lagged polling of the device stop flags.
struct pinned_flags {
    gkomi_stop_flags* host{};
    explicit pinned_flags(size_type n)
    {
        void* p = nullptr;
        GKO_CALL(gkomi_host_malloc(n * sizeof(gkomi_stop_flags), &p));
        host = static_cast<gkomi_stop_flags*>(p);
        std::memset(host, 0, n * sizeof(gkomi_stop_flags));
    }
    ~pinned_flags() { gkomi_host_free(host); }
    pinned_flags(const pinned_flags&) = delete;
    pinned_flags& operator=(const pinned_flags&) = delete;
};

struct event {
    int device;
    void* ev{};
    explicit event(int dev) : device{dev} { GKO_CALL(gkomi_event_create(dev, &ev)); }
    ~event() { gkomi_event_destroy(device, ev); }
    event(const event&) = delete;
    event& operator=(const event&) = delete;
};

bool needs_tau(const gkomi_criteria& c)
{
    for (int k = 0; k < c.num; ++k)
        if (c.c[k].kind == GKOMI_CRIT_RESIDUAL_NORM) return true;
    return false;
}

// Baselines of the residual criteria (core/stop/residual_norm.cpp:93-162):
// rhs_norm -> ||b||, initial_resnorm -> ||r0||, absolute -> 1; one row per criterion.
template <typename V>
std::unique_ptr<Vec<V>> baselines(std::shared_ptr<const Executor> exec, const gkomi_criteria& c,
                                  const Vec<V>* b, const Vec<V>* r0)
{
    const auto k = b->get_size().cols;
    auto orig = Vec<V>::create(exec, dim<2>{static_cast<size_type>(std::max(c.num, 1)), k});
    orig->fill(V{1});
    for (int i = 0; i < c.num; ++i) {
        if (c.c[i].kind == GKOMI_CRIT_ITERATION || c.c[i].baseline == GKOMI_BASELINE_ABSOLUTE) {
            continue;
        }
        auto row = Vec<V>::create_view(exec, dim<2>{1, k}, orig->get_values() + i * k, k);
        (c.c[i].baseline == GKOMI_BASELINE_RHS_NORM ? b : r0)->compute_norm2(row.get());
    }
    return orig;
}

template <typename V>
std::vector<double> host_norms(const Vec<V>* v)
{
    auto exec = v->get_executor();
    auto n = Vec<V>::create(exec, dim<2>{1, v->get_size().cols});
    v->compute_norm2(n.get());
    auto h = n->to_host();
    return std::vector<double>(h.begin(), h.end());
}

std::vector<stopping_status> host_status(const array<uint8>& st)
{
    auto raw = st.to_host();
    std::vector<stopping_status> out(raw.size());
    for (size_type i = 0; i < raw.size(); ++i) {
        std::memcpy(&out[i], &raw[i], 1);
    }
    return out;
}

// device-resident flags with `iteration` set from the host
void set_iteration(const std::shared_ptr<const Executor>& exec, gkomi_stop_flags* dflags,
                   int64 iter)
{
    gkomi_stop_flags f{};
    f.iteration = iter;
    exec->copy_from(exec->get_master().get(), sizeof(f), reinterpret_cast<const char*>(&f),
                    reinterpret_cast<char*>(dflags));
}

gkomi_stop_flags read_flags(const std::shared_ptr<const Executor>& exec,
                            const gkomi_stop_flags* dflags)
{
    return exec->copy_val_to_host(dflags);
}

// Lagged polling of a fused solve: `iterate()` enqueues one chunk of
// iterations; two chunks stay in flight, and the host waits only for the
// older chunk's copy of the device flags before deciding to go on.
template <typename Iterate>
void run_until_stopped(const gkomi_ctx* ctx, const gkomi_stop_flags* dflags, Iterate iterate)
{
    pinned_flags slots(2);
    event e0(ctx->device), e1(ctx->device);
    event* evs[2] = {&e0, &e1};
    auto enqueue = [&](int slot) {
        iterate();
        GKO_CALL(gkomi_memcpy_async(ctx, &slots.host[slot], dflags, sizeof(gkomi_stop_flags)));
        GKO_CALL(gkomi_event_record(ctx, evs[slot]->ev));
    };
    enqueue(0);
    for (int64 i = 0;; ++i) {
        enqueue(static_cast<int>((i + 1) % 2));
        GKO_CALL(gkomi_event_synchronize(evs[i % 2]->ev));
        if (slots.host[i % 2].all_stopped) break;
    }
}

}  // namespace

// ------------------------------------------------------------- Jacobi ----

namespace preconditioner {

template <typename V, typename I>
Jacobi<V, I>::Jacobi(std::shared_ptr<const Executor> exec, parameters_type p, const LinOp* a)
    : LinOp(exec, a->get_size()), inv_diag_(exec, a->get_size().rows)
{
    if (p.max_block_size != 1) {
        GKO_NOT_SUPPORTED_MSG(
            "only scalar Jacobi (with_max_block_size(1u)) is part of this path");
    }
    if (a->get_size().rows != a->get_size().cols) {
        throw DimensionMismatch(__FILE__, __LINE__, "Jacobi: system matrix is not square");
    }
    std::unique_ptr<matrix::Csr<V, I>> converted;
    const matrix::Csr<V, I>* csr = dynamic_cast<const matrix::Csr<V, I>*>(a);
    if (!csr) {
        using Other = std::conditional_t<std::is_same<V, double>::value, float, double>;
        if (auto o = dynamic_cast<const matrix::Csr<Other, I>*>(a)) {
            converted = o->template convert_precision<V>();
            csr = converted.get();
        } else {
            GKO_NOT_SUPPORTED_MSG("Jacobi: the system matrix must be Csr");
        }
    }
    const auto d = csr->descriptor();
    GKO_CALL(sk<V>::jacobi_generate(ctx_of(exec), &d, inv_diag_.get_data()));
}

template <typename V, typename I>
void Jacobi<V, I>::apply_impl(const LinOp* b, LinOp* x) const
{
    precision_dispatch<V>(
        [this](auto db, auto dx) {
            GKO_CALL(sk<V>::jacobi_apply(ctx_of(this->get_executor()), db->get_size().rows,
                                         db->get_size().cols, inv_diag_.get_const_data(),
                                         db->get_const_values(), db->get_stride(),
                                         dx->get_values(), dx->get_stride()));
        },
        b, x);
}

template <typename V, typename I>
void Jacobi<V, I>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                              LinOp* x) const
{
    precision_dispatch<V>(
        [this](auto da, auto db, auto dbeta, auto dx) {
            GKO_CALL(sk<V>::jacobi_adv(ctx_of(this->get_executor()), db->get_size().rows,
                                       db->get_size().cols, inv_diag_.get_const_data(),
                                       da->get_const_values(), db->get_const_values(),
                                       db->get_stride(), dbeta->get_const_values(),
                                       dx->get_values(), dx->get_stride()));
        },
        alpha, b, beta, x);
}

template class Jacobi<double, int32>;
template class Jacobi<float, int32>;

}  // namespace preconditioner

namespace solver {

// --------------------------------------------------------- KrylovBase ----

template <typename V>
KrylovBase<V>::KrylovBase(std::shared_ptr<const Executor> exec, const iterative_parameters& p,
                          std::shared_ptr<const LinOp> a)
    : LinOp(exec, a ? a->get_size() : dim<2>{}), system_matrix_{std::move(a)}
{
    if (system_matrix_ && system_matrix_->get_size().rows != system_matrix_->get_size().cols) {
        throw DimensionMismatch(__FILE__, __LINE__, "solver: system matrix is not square");
    }
    crit_ = p.criteria.resolve(exec);
    if (p.generated_preconditioner) {
        precond_ = p.generated_preconditioner;
    } else if (p.preconditioner) {
        precond_ = p.preconditioner.on(exec)->generate(system_matrix_);
    }
    if (auto j = dynamic_cast<const preconditioner::Jacobi<V, int32>*>(precond_.get())) {
        inv_diag_ = j->get_inverse_diagonal();
    }
}

// Solver::apply_impl(alpha, b, beta, x) (core/solver/cg.cpp:199-214):
// x = alpha * solve(b; x) + beta * x
template <typename V>
void KrylovBase<V>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                               LinOp* x) const
{
    precision_dispatch<V>(
        [this](auto da, auto db, auto dbeta, auto dx) {
            auto xc = dx->clone();
            this->apply_impl(db, xc.get());
            dx->scale(dbeta);
            dx->add_scaled(da, xc.get());
        },
        alpha, b, beta, x);
}

template <typename V>
void KrylovBase<V>::log_result(const Vec<V>* b, const Vec<V>* x) const
{
    if (this->get_loggers().empty()) return;
    auto r = Vec<V>::create(this->get_executor(), b->get_size());
    r->copy_from(b);
    auto one = scalar<V>(this->get_executor(), V{1});
    auto neg = scalar<V>(this->get_executor(), V{-1});
    system_matrix_->apply(neg.get(), x, one.get(), r.get());
    const auto norms = host_norms(r.get());
    for (auto& l : this->get_loggers()) {
        l->on_iteration_complete(this, stats_.iterations, norms, stats_.status, true);
    }
}

// ------------------------------------------------------------------ Cg ----

template <typename V>
void Cg<V>::apply_impl(const LinOp* b, LinOp* x) const
{
    precision_dispatch<V>([this](auto db, auto dx) { this->apply_dense(db, dx); }, b, x);
}

template <typename V>
bool Cg<V>::try_fused(const Vec<V>* b, Vec<V>* x) const
{
    auto csr = dynamic_cast<const matrix::Csr<V, int32>*>(this->system_matrix_.get());
    if (!csr || b->get_size().cols != 1 || b->get_stride() != 1 || x->get_stride() != 1) {
        return false;
    }
    if (csr->get_effective_strategy() != "classical") return false;
    if (this->precond_ && !this->inv_diag_) return false;
    auto exec = this->get_executor();
    const size_type n = b->get_size().rows;
    array<V> r(exec, n), z(exec, n), pa(exec, n), pb(exec, n), q(exec, n);
    array<V> scalars(exec, 8), orig(exec, GKOMI_MAX_CRITERIA);
    array<uint8> status(exec, 1);
    array<char> flags(exec, sizeof(gkomi_stop_flags));
    const size_type wb = gkomi_cg_work_bytes(n);
    array<char> work(exec, wb);
    gkomi_cg_state st{};
    st.n = static_cast<int64_t>(n);
    st.x = x->get_values();
    st.b = b->get_const_values();
    st.r = r.get_data();
    st.z = z.get_data();
    st.p_a = pa.get_data();
    st.p_b = pb.get_data();
    st.q = q.get_data();
    st.inv_diag = this->inv_diag_;
    st.scalars = scalars.get_data();
    st.orig_tau = orig.get_data();
    st.status = status.get_data();
    st.flags = reinterpret_cast<gkomi_stop_flags*>(flags.get_data());
    st.work = work.get_data();
    const auto d = csr->descriptor();
    const gkomi_ctx* ctx = ctx_of(exec);
    GKO_CALL(sk<V>::cg_fused_start(ctx, &d, &this->crit_, &st));
    // two chunks in flight: poll chunk i while chunk i + 1 runs
    pinned_flags slots(2);
    const int dev = ctx->device;
    event e0(dev), e1(dev);
    event* evs[2] = {&e0, &e1};
    int64 issued = 0;
    auto enqueue = [&](int slot) {
        GKO_CALL(sk<V>::cg_fused_iterate(ctx, &d, &this->crit_, &st, this->chunk_));
        GKO_CALL(gkomi_memcpy_async(ctx, &slots.host[slot], st.flags, sizeof(gkomi_stop_flags)));
        GKO_CALL(gkomi_event_record(ctx, evs[slot]->ev));
        ++issued;
    };
    // the start kernels may already have stopped the solve (iteration 0)
    GKO_CALL(gkomi_memcpy_async(ctx, &slots.host[0], st.flags, sizeof(gkomi_stop_flags)));
    exec->synchronize();
    bool done = slots.host[0].all_stopped != 0;
    if (!done) {
        enqueue(0);
        for (int64 i = 0;; ++i) {
            enqueue(static_cast<int>((i + 1) % 2));
            GKO_CALL(gkomi_event_synchronize(evs[i % 2]->ev));
            if (slots.host[i % 2].all_stopped) break;
        }
    }
    exec->synchronize();
    const auto f = read_flags(exec, st.flags);
    this->stats_.iterations = f.iteration;
    this->stats_.status = host_status(status);
    this->stats_.used_fused_path = true;
    return true;
}

// Several right-hand sides: the fused k-column CG (four kernels per
// iteration, per-column scalars and status), with the same chunked lagged
// polling.  Any Csr strategy works (one lane group per row).
template <typename V>
bool Cg<V>::try_fused_block(const Vec<V>* b, Vec<V>* x) const
{
    auto csr = dynamic_cast<const matrix::Csr<V, int32>*>(this->system_matrix_.get());
    const size_type k = b->get_size().cols;
    if (!csr || k < 2) return false;
    if (this->precond_ && !this->inv_diag_) return false;
    auto exec = this->get_executor();
    const size_type n = b->get_size().rows;
    array<V> r(exec, n * k), pa(exec, n * k), pb(exec, n * k), q(exec, n * k);
    array<V> z(exec, this->inv_diag_ ? n * k : 0);
    array<V> scalars(exec, 5 * k), orig(exec, GKOMI_MAX_CRITERIA * k);
    array<uint8> status(exec, k);
    array<char> flags(exec, sizeof(gkomi_stop_flags));
    array<char> work(exec, gkomi_block_cg_work_bytes(n, k));
    gkomi_block_cg_state st{};
    st.n = static_cast<int64_t>(n);
    st.nrhs = static_cast<int64_t>(k);
    st.x = x->get_values();
    st.x_stride = static_cast<int64_t>(x->get_stride());
    st.b = b->get_const_values();
    st.b_stride = static_cast<int64_t>(b->get_stride());
    st.r = r.get_data();
    st.z = this->inv_diag_ ? z.get_data() : r.get_data();
    st.p_a = pa.get_data();
    st.p_b = pb.get_data();
    st.q = q.get_data();
    st.inv_diag = this->inv_diag_;
    st.scalars = scalars.get_data();
    st.orig_tau = orig.get_data();
    st.status = status.get_data();
    st.flags = reinterpret_cast<gkomi_stop_flags*>(flags.get_data());
    st.work = work.get_data();
    const auto d = csr->descriptor();
    const gkomi_ctx* ctx = ctx_of(exec);
    GKO_CALL(sk<V>::block_cg_start(ctx, &d, &this->crit_, &st));
    pinned_flags slots(2);
    const int dev = ctx->device;
    event e0(dev), e1(dev);
    event* evs[2] = {&e0, &e1};
    auto enqueue = [&](int slot) {
        GKO_CALL(sk<V>::block_cg_iterate(ctx, &d, &this->crit_, &st, this->chunk_));
        GKO_CALL(gkomi_memcpy_async(ctx, &slots.host[slot], st.flags, sizeof(gkomi_stop_flags)));
        GKO_CALL(gkomi_event_record(ctx, evs[slot]->ev));
    };
    enqueue(0);
    for (int64 i = 0;; ++i) {
        enqueue(static_cast<int>((i + 1) % 2));
        GKO_CALL(gkomi_event_synchronize(evs[i % 2]->ev));
        if (slots.host[i % 2].all_stopped) break;
    }
    exec->synchronize();
    const auto f = read_flags(exec, st.flags);
    this->stats_.iterations = f.iteration;
    this->stats_.status = host_status(status);
    this->stats_.used_fused_path = true;
    return true;
}

// Generic path: core/solver/cg.cpp:93-196, with the check evaluated on the
// device and its result read back once per iteration.
template <typename V>
void Cg<V>::apply_dense(const Vec<V>* b, Vec<V>* x) const
{
    if (!this->system_matrix_) return;
    if (!try_fused(b, x) && !try_fused_block(b, x)) {
        auto exec = this->get_executor();
        const auto ctx = ctx_of(exec);
        const size_type n = b->get_size().rows, k = b->get_size().cols;
        auto r = Vec<V>::create(exec, dim<2>{n, k});
        auto z = Vec<V>::create(exec, dim<2>{n, k});
        auto p = Vec<V>::create(exec, dim<2>{n, k});
        auto q = Vec<V>::create(exec, dim<2>{n, k});
        auto beta = Vec<V>::create(exec, dim<2>{1, k});
        auto prev_rho = Vec<V>::create(exec, dim<2>{1, k});
        auto rho = Vec<V>::create(exec, dim<2>{1, k});
        auto tau = Vec<V>::create(exec, dim<2>{1, k});
        array<uint8> status(exec, k);
        array<char> flags(exec, sizeof(gkomi_stop_flags));
        auto dflags = reinterpret_cast<gkomi_stop_flags*>(flags.get_data());
        auto one = scalar<V>(exec, V{1});
        auto neg = scalar<V>(exec, V{-1});
        GKO_CALL(sk<V>::cg_init(ctx, n, k, b->get_const_values(), b->get_stride(),
                                r->get_values(), z->get_values(), p->get_values(),
                                q->get_values(), k, prev_rho->get_values(), rho->get_values(),
                                status.get_data()));
        this->system_matrix_->apply(neg.get(), x, one.get(), r.get());
        auto orig = baselines<V>(exec, this->crit_, b, r.get());
        const bool want_tau = needs_tau(this->crit_);
        int64 iter = -1;
        while (true) {
            if (this->precond_) {
                this->precond_->apply(r.get(), z.get());
            } else {
                z->copy_from(r.get());
            }
            r->compute_dot(z.get(), rho.get());
            ++iter;
            set_iteration(exec, dflags, iter);
            if (want_tau) r->compute_norm2(tau.get());
            GKO_CALL(sk<V>::check(ctx, &this->crit_, k, want_tau ? tau->get_values() : nullptr,
                                  rho->get_values(), orig->get_values(), 1, status.get_data(),
                                  dflags));
            if (read_flags(exec, dflags).all_stopped) break;
            GKO_CALL(sk<V>::cg_step_1(ctx, n, k, p->get_values(), z->get_values(), k,
                                      rho->get_values(), prev_rho->get_values(),
                                      status.get_data()));
            this->system_matrix_->apply(p.get(), q.get());
            p->compute_dot(q.get(), beta.get());
            GKO_CALL(sk<V>::cg_step_2(ctx, n, k, x->get_values(), x->get_stride(),
                                      r->get_values(), p->get_values(), q->get_values(), k,
                                      beta->get_values(), rho->get_values(),
                                      status.get_data()));
            std::swap(prev_rho, rho);
        }
        this->stats_.iterations = iter;
        this->stats_.status = host_status(status);
        this->stats_.used_fused_path = false;
    }
    this->log_result(b, x);
}

// ------------------------------------------------------------ Bicgstab ----

template <typename V>
void Bicgstab<V>::apply_impl(const LinOp* b, LinOp* x) const
{
    precision_dispatch<V>([this](auto db, auto dx) { this->apply_dense(db, dx); }, b, x);
}

// One right-hand side over a classical Csr with the identity or scalar
// Jacobi: the fused iteration (five kernels, checks on the device).
template <typename V>
bool Bicgstab<V>::try_fused(const Vec<V>* b, Vec<V>* x) const
{
    auto csr = dynamic_cast<const matrix::Csr<V, int32>*>(this->system_matrix_.get());
    if (!csr || b->get_size().cols != 1 || b->get_stride() != 1 || x->get_stride() != 1) {
        return false;
    }
    if (csr->get_effective_strategy() != "classical") return false;
    if (this->precond_ && !this->inv_diag_) return false;
    auto exec = this->get_executor();
    const size_type n = b->get_size().rows;
    const size_type nm = this->inv_diag_ ? n : 0;
    array<V> r(exec, n), rr(exec, n), p(exec, n), v(exec, n), s(exec, n), t(exec, n);
    array<V> y(exec, nm), z(exec, nm);
    array<V> scalars(exec, 8), orig(exec, GKOMI_MAX_CRITERIA);
    array<uint8> status(exec, 1);
    array<char> flags(exec, sizeof(gkomi_stop_flags));
    array<char> work(exec, gkomi_bicgstab_work_bytes(static_cast<int64_t>(n)));
    gkomi_bicgstab_state st{};
    st.n = static_cast<int64_t>(n);
    st.x = x->get_values();
    st.b = b->get_const_values();
    st.r = r.get_data();
    st.rr = rr.get_data();
    st.p = p.get_data();
    st.v = v.get_data();
    st.s = s.get_data();
    st.t = t.get_data();
    st.y = nm ? y.get_data() : nullptr;
    st.z = nm ? z.get_data() : nullptr;
    st.inv_diag = this->inv_diag_;
    st.scalars = scalars.get_data();
    st.orig_tau = orig.get_data();
    st.status = status.get_data();
    st.flags = reinterpret_cast<gkomi_stop_flags*>(flags.get_data());
    st.work = work.get_data();
    const auto d = csr->descriptor();
    const gkomi_ctx* ctx = ctx_of(exec);
    GKO_CALL(sk<V>::bicg_fused_start(ctx, &d, &this->crit_, &st));
    run_until_stopped(ctx, st.flags, [&] {
        GKO_CALL(sk<V>::bicg_fused_iterate(ctx, &d, &this->crit_, &st, this->chunk_));
    });
    exec->synchronize();
    const auto f = read_flags(exec, st.flags);
    this->stats_.iterations = f.iteration;
    this->stats_.status = host_status(status);
    this->stats_.used_fused_path = true;
    return true;
}

// core/solver/bicgstab.cpp:95-255 with device-side checks.
template <typename V>
void Bicgstab<V>::apply_dense(const Vec<V>* b, Vec<V>* x) const
{
    if (!this->system_matrix_) return;
    if (try_fused(b, x)) {
        this->log_result(b, x);
        return;
    }
    auto exec = this->get_executor();
    const auto ctx = ctx_of(exec);
    const size_type n = b->get_size().rows, k = b->get_size().cols;
    auto mk = [&] { return Vec<V>::create(exec, dim<2>{n, k}); };
    auto sc = [&] { return Vec<V>::create(exec, dim<2>{1, k}); };
    auto r = mk(), z = mk(), y = mk(), v = mk(), s = mk(), t = mk(), p = mk(), rr = mk();
    auto alpha = sc(), beta = sc(), gamma = sc(), prev_rho = sc(), rho = sc(), omega = sc(),
         tau = sc();
    array<uint8> status(exec, k);
    array<char> flags(exec, sizeof(gkomi_stop_flags));
    auto dflags = reinterpret_cast<gkomi_stop_flags*>(flags.get_data());
    auto one = scalar<V>(exec, V{1});
    auto neg = scalar<V>(exec, V{-1});
    GKO_CALL(sk<V>::bicg_init(ctx, n, k, b->get_const_values(), b->get_stride(), r->get_values(),
                              rr->get_values(), y->get_values(), s->get_values(),
                              t->get_values(), z->get_values(), v->get_values(), p->get_values(),
                              k, prev_rho->get_values(), rho->get_values(), alpha->get_values(),
                              beta->get_values(), gamma->get_values(), omega->get_values(),
                              status.get_data()));
    this->system_matrix_->apply(neg.get(), x, one.get(), r.get());
    auto orig = baselines<V>(exec, this->crit_, b, r.get());
    rr->copy_from(r.get());
    const bool want_tau = needs_tau(this->crit_);
    auto prec = [&](const Vec<V>* in, Vec<V>* out) {
        if (this->precond_) {
            this->precond_->apply(in, out);
        } else {
            out->copy_from(in);
        }
    };
    int64 iter = -1;
    while (true) {
        ++iter;
        rr->compute_dot(r.get(), rho.get());
        set_iteration(exec, dflags, iter);
        if (want_tau) r->compute_norm2(tau.get());
        GKO_CALL(sk<V>::check(ctx, &this->crit_, k, want_tau ? tau->get_values() : nullptr,
                              rho->get_values(), orig->get_values(), 1, status.get_data(),
                              dflags));
        if (read_flags(exec, dflags).all_stopped) break;
        GKO_CALL(sk<V>::bicg_step_1(ctx, n, k, r->get_values(), p->get_values(),
                                    v->get_values(), k, rho->get_values(),
                                    prev_rho->get_values(), alpha->get_values(),
                                    omega->get_values(), status.get_data()));
        prec(p.get(), y.get());
        this->system_matrix_->apply(y.get(), v.get());
        rr->compute_dot(v.get(), beta.get());
        GKO_CALL(sk<V>::bicg_step_2(ctx, n, k, r->get_values(), s->get_values(), v->get_values(),
                                    k, rho->get_values(), alpha->get_values(),
                                    beta->get_values(), status.get_data()));
        if (want_tau) s->compute_norm2(tau.get());
        GKO_CALL(sk<V>::check(ctx, &this->crit_, k, want_tau ? tau->get_values() : nullptr,
                              rho->get_values(), orig->get_values(), 0, status.get_data(),
                              dflags));
        const auto f = read_flags(exec, dflags);
        if (f.one_changed) {
            GKO_CALL(sk<V>::bicg_finalize(ctx, n, k, x->get_values(), x->get_stride(),
                                          y->get_values(), k, alpha->get_values(),
                                          status.get_data()));
        }
        if (f.all_stopped) break;
        prec(s.get(), z.get());
        this->system_matrix_->apply(z.get(), t.get());
        s->compute_dot(t.get(), gamma.get());
        t->compute_dot(t.get(), beta.get());
        GKO_CALL(sk<V>::bicg_step_3(ctx, n, k, x->get_values(), x->get_stride(),
                                    r->get_values(), s->get_values(), t->get_values(),
                                    y->get_values(), z->get_values(), k, alpha->get_values(),
                                    beta->get_values(), gamma->get_values(),
                                    omega->get_values(), status.get_data()));
        std::swap(prev_rho, rho);
    }
    this->stats_.iterations = iter;
    this->stats_.status = host_status(status);
    this->stats_.used_fused_path = false;
    this->log_result(b, x);
}

// ------------------------------------------------------------------ Ir ----

template <typename V>
Ir<V>::Ir(std::shared_ptr<const Executor> exec, const parameters_type& p,
          std::shared_ptr<const LinOp> a)
    : LinOp(exec, a ? a->get_size() : dim<2>{}),
      system_matrix_{std::move(a)},
      relaxation_{p.relaxation_factor}
{
    crit_ = p.criteria.resolve(exec);
    if (p.generated_solver) {
        inner_ = p.generated_solver;
    } else if (p.solver) {
        auto factory = p.solver.on(exec);
        // an inner solver of the other precision gets the converted matrix
        using Other = std::conditional_t<std::is_same<V, double>::value, float, double>;
        std::shared_ptr<const LinOp> inner_a = system_matrix_;
        auto probe = factory.get();
        const bool other_cg = dynamic_cast<const typename Cg<Other>::Factory*>(probe) ||
                              dynamic_cast<const typename Bicgstab<Other>::Factory*>(probe);
        if (other_cg) {
            if (auto csr = dynamic_cast<const matrix::Csr<V, int32>*>(system_matrix_.get())) {
                // the inner copy is the library's own: it also gets the 2-byte
                // column offsets when they fit (6 instead of 8 bytes per FP32
                // entry, bit-identical results)
                auto lower = csr->template convert_precision<Other>();
                lower->compact_indices();
                // rows of nearly equal length also get the sliced copy
                // (SELL-P layout: coalesced loads, a lane per row)
                const auto n = lower->get_size().rows;
                const double mean =
                    n ? double(lower->get_num_stored_elements()) / double(n) : 0.0;
                if (n > 0 && double(lower->max_row_nnz()) <= 1.25 * mean + 4.0) {
                    lower->build_slices(true);
                }
                inner_a = std::shared_ptr<const LinOp>(std::move(lower));
            }
        }
        inner_ = factory->generate(inner_a);
    }
}

template <typename V>
void Ir<V>::apply_impl(const LinOp* b, LinOp* x) const
{
    precision_dispatch<V>(
        [this](auto db, auto dx) {
            if (!system_matrix_) return;
            auto exec = this->get_executor();
            const auto ctx = ctx_of(exec);
            const size_type n = db->get_size().rows, k = db->get_size().cols;
            auto residual = Vec<V>::create(exec, dim<2>{n, k});
            auto inner_solution = Vec<V>::create(exec, dim<2>{n, k});
            auto tau = Vec<V>::create(exec, dim<2>{1, k});
            auto one = scalar<V>(exec, V{1});
            auto neg = scalar<V>(exec, V{-1});
            auto relax = scalar<V>(exec, static_cast<V>(relaxation_));
            array<uint8> status(exec, k);
            status.fill(0);
            array<char> flags(exec, sizeof(gkomi_stop_flags));
            auto dflags = reinterpret_cast<gkomi_stop_flags*>(flags.get_data());
            residual->copy_from(db);
            system_matrix_->apply(neg.get(), dx, one.get(), residual.get());
            auto orig = baselines<V>(exec, crit_, db, residual.get());
            int64 iter = -1;
            stats_.inner_iterations.clear();
            while (true) {
                ++iter;
                set_iteration(exec, dflags, iter);
                bool all_stopped;
                if (iter == 0) {
                    residual->compute_norm2(tau.get());
                    GKO_CALL(sk<V>::check(ctx, &crit_, k, tau->get_values(), nullptr,
                                          orig->get_values(), 1, status.get_data(), dflags));
                    all_stopped = read_flags(exec, dflags).all_stopped;
                } else {
                    // update_residual (core/solver/update_residual.hpp): the
                    // residual check is skipped first, then r = b - A x
                    GKO_CALL(sk<V>::check(ctx, &crit_, k, nullptr, nullptr, orig->get_values(), 0,
                                          status.get_data(), dflags));
                    all_stopped = read_flags(exec, dflags).all_stopped;
                    if (!all_stopped) {
                        residual->copy_from(db);
                        system_matrix_->apply(neg.get(), dx, one.get(), residual.get());
                        residual->compute_norm2(tau.get());
                        GKO_CALL(sk<V>::check(ctx, &crit_, k, tau->get_values(), nullptr,
                                              orig->get_values(), 1, status.get_data(), dflags));
                        all_stopped = read_flags(exec, dflags).all_stopped;
                    }
                }
                if (all_stopped) break;
                if (inner_) {
                    inner_solution->copy_from(residual.get());
                    inner_->apply(residual.get(), inner_solution.get());
                    dx->add_scaled(relax.get(), inner_solution.get());
                    if (auto k = dynamic_cast<const KrylovBase<double>*>(inner_.get())) {
                        stats_.inner_iterations.push_back(k->get_last_stats().iterations);
                    } else if (auto k = dynamic_cast<const KrylovBase<float>*>(inner_.get())) {
                        stats_.inner_iterations.push_back(k->get_last_stats().iterations);
                    }
                } else {
                    dx->add_scaled(relax.get(), residual.get());
                }
            }
            stats_.iterations = iter;
            stats_.status = host_status(status);
            if (!this->get_loggers().empty()) {
                residual->copy_from(db);
                system_matrix_->apply(neg.get(), dx, one.get(), residual.get());
                const auto norms = host_norms(residual.get());
                for (auto& l : this->get_loggers()) {
                    l->on_iteration_complete(this, iter, norms, stats_.status, true);
                }
            }
        },
        b, x);
}

template <typename V>
void Ir<V>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta, LinOp* x) const
{
    precision_dispatch<V>(
        [this](auto da, auto db, auto dbeta, auto dx) {
            auto xc = dx->clone();
            this->apply_impl(db, xc.get());
            dx->scale(dbeta);
            dx->add_scaled(da, xc.get());
        },
        alpha, b, beta, x);
}

template class KrylovBase<double>;
template class KrylovBase<float>;
template class Cg<double>;
template class Cg<float>;
template class Bicgstab<double>;
template class Bicgstab<float>;
template class Ir<double>;
template class Ir<float>;

}  // namespace solver
}  // namespace gko

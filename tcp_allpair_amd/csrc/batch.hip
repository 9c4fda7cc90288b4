// Batches of independent replicas (tcp_batch_* in include/tcp_gpu.h).
//
// A batch owns one context per replica (all state, histograms and the
// reciprocal-space work stay per replica) and steps them together: every
// O(N) kernel of the step runs once for all replicas with blockIdx.y as the
// replica index, the pair work of all replicas goes through ONE persistent
// launch that takes (replica, block pair) items from a device queue, and the
// whole step is captured once in a hipGraph and replayed.  The kernel bodies
// are those of the single-context path with the same block decomposition, so
// every replica reproduces its standalone run bit for bit: always in FP64, and
// in mixed precision at the standalone run's tile shape (4 x 4 tiles below
// about 46000 particles, 8 x 8 above).  The autotuner may pick another shape
// for the persistent mixed kernel (a different, equally valid FP32 grouping).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "tcp_ctx.h"

using namespace tcp;

namespace {

constexpr double kPi = 3.1415926535897931;  // data.f90:23

thread_local std::string g_batch_create_error;

}  // namespace

struct tcp_batch {
    int nrep = 0;
    int precision = TCP_FP64;
    int device = 0;
    int cus = 1;
    std::vector<tcp_ctx*> reps;
    hipStream_t stream = nullptr;
    int maxblk = 1;           // O(N) blocks of the largest replica
    int maxn = 2;             // particles of the largest replica
    int recip = TCP_RECIP_NONE;
    int kmax = 0;
    SpmeBatchPlan spme;       // TCP_RECIP_SPME: launch sizes and the FFT pass tables
    bool kelbg = false;
    // device tables
    std::vector<BatchRep> h_reps;
    BatchRep* d_reps = nullptr;
    std::vector<double*> pf, pu;   // per replica pair partials, slots for every shape
    PairItem* d_items = nullptr;
    int item_cap = 0, nitems = 0;
    int nfine = 0;            // FP64: the last nfine items belong to 4 x 4 tile replicas
    int* d_queue = nullptr;
    uint64_t* d_counter = nullptr;  // [0] noise step counter, [1] trace row
    unsigned long long* d_rdf_scratch = nullptr;
    double* d_trace = nullptr;
    size_t trace_cap = 0;  // trace rows (steps x replicas)
    // persistent launch shape
    int shape = 0;
    int grid = 1;
    int grid_fine = 1;        // FP64: resident blocks of the 4 x 4 tile launch
    // captured step
    hipGraphExec_t exec = nullptr;
    struct Key {
        int thermo = -1;
        double dt = 0, tau = 0, fric = 0;
        uint64_t seed = 0;
        double* trace = nullptr;
        int shape = -1, grid = -1, nitems = -1, nfine = -1, grid_fine = -1;
        bool operator==(const Key& o) const
        {
            return thermo == o.thermo && dt == o.dt && tau == o.tau && fric == o.fric &&
                   seed == o.seed && trace == o.trace && shape == o.shape && grid == o.grid &&
                   nitems == o.nitems && nfine == o.nfine && grid_fine == o.grid_fine;
        }
    } key;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // A second stream (same priority), forked from and joined back to
    // `stream` by these events: FP64 runs its 4 x 4 tile launch there, so its
    // blocks fill the CUs the tail of the 8 x 8 tile launch leaves idle; mixed
    // runs SPME there beside the pair launch
    hipStream_t stream2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    double pair_ms = 0, steps_ms = 0;
    uint64_t counter = 0;
    std::vector<void*> allocs;
    std::string err;
};

namespace {

int bfail(tcp_batch* b, int code, const std::string& msg)
{
    if (b) b->err = msg;
    else g_batch_create_error = msg;
    return code;
}

#define BHIP(b, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bfail(b, TCP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

#define BTRY(expr)             \
    do {                       \
        int rc_ = (expr);      \
        if (rc_) return rc_;   \
    } while (0)

template <typename T>
int balloc(tcp_batch* b, T** p, size_t count)
{
    void* q = nullptr;
    BHIP(b, hipMalloc(&q, count * sizeof(T) + 64));
    b->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

// Grows a batch buffer to at least count elements, freeing the old one once
// the stream has drained (contents are not kept).  The captured step refers to
// the old address, so it is dropped and captured again on the next steps.
template <typename T>
int bgrow(tcp_batch* b, T** p, size_t count)
{
    if (*p) {
        BHIP(b, hipStreamSynchronize(b->stream));
        if (b->exec) {
            BHIP(b, hipGraphExecDestroy(b->exec));
            b->exec = nullptr;
        }
        auto it = std::find(b->allocs.begin(), b->allocs.end(), (void*)*p);
        if (it != b->allocs.end()) b->allocs.erase(it);
        BHIP(b, hipFree(*p));
        *p = nullptr;
    }
    return balloc(b, p, count);
}

bool mixed(const tcp_batch* b) { return b->precision == TCP_MIXED; }

// The block shape of a replica's pair work: the batch's in mixed precision;
// in FP64 its own context's (8 x 8 or, for small systems, 4 x 4 tiles), so that
// every replica keeps its standalone summation order.
int rep_shape(const tcp_batch* b, const tcp_ctx* c, int shape)
{
    return mixed(b) ? shape : c->n3_shape;
}

int slots_of(const tcp_batch* b, const tcp_ctx* c, int shape)
{
    return n3_block_rows_shape(c->N, rep_shape(b, c, shape));
}

// Work list of the persistent pair launch, largest items first.  FP64: the
// items of 8 x 8 tile replicas, then (*nfine of them) those of 4 x 4 tile
// replicas, each group sorted on its own (one launch per group).
std::vector<PairItem> make_items(const tcp_batch* b, int shape, int* nfine)
{
    std::vector<PairItem> items[2];
    for (int r = 0; r < b->nrep; ++r) {
        const tcp_ctx* c = b->reps[r];
        const int rs = rep_shape(b, c, shape);
        const int bt = n3_shape(rs).tiles;
        const int ntiles = (c->N + 63) / 64;
        const int rows = n3_block_rows_shape(c->N, rs);
        auto& out = items[!mixed(b) && rs == kFineN3Shape ? 1 : 0];
        for (int br = 0; br < rows; ++br) {
            const int ti = std::min(bt, ntiles - br * bt);
            for (int bc = br; bc < rows; ++bc) {
                const int tj = std::min(bt, ntiles - bc * bt);
                const int cost = bc == br ? ti * (ti + 1) / 2 : ti * tj;
                out.push_back(PairItem{r, br, bc, cost});
            }
        }
    }
    for (auto& v : items)
        std::stable_sort(v.begin(), v.end(),
                         [](const PairItem& x, const PairItem& y) { return x.cost > y.cost; });
    *nfine = (int)items[1].size();
    items[0].insert(items[0].end(), items[1].begin(), items[1].end());
    return items[0];
}

// Uploads the work list and the replica table for a block shape.
int set_shape(tcp_batch* b, int shape)
{
    int nfine = 0;
    const std::vector<PairItem> items = make_items(b, shape, &nfine);
    if ((int)items.size() > b->item_cap) {
        BTRY(bgrow(b, &b->d_items, items.size()));
        b->item_cap = (int)items.size();
    }
    BHIP(b, hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(PairItem),
                           hipMemcpyHostToDevice, b->stream));
    b->nitems = (int)items.size();
    b->nfine = nfine;
    for (int r = 0; r < b->nrep; ++r) {
        b->h_reps[r].slots = slots_of(b, b->reps[r], shape);
        b->h_reps[r].pu_row = n3_shape(rep_shape(b, b->reps[r], shape)).tiles * 64;
    }
    BHIP(b, hipMemcpyAsync(b->d_reps, b->h_reps.data(), b->nrep * sizeof(BatchRep),
                           hipMemcpyHostToDevice, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->shape = shape;
    b->grid = b->cus * pair_batch_occupancy(mixed(b), mixed(b) ? shape : 0, true, b->kelbg);
    b->grid_fine = b->cus * pair_batch_occupancy(false, kFineN3Shape, true, b->kelbg);
    return 0;
}

void launch_pair(tcp_batch* b)
{
    if (mixed(b))
        launch_pair_mixed_n3_batch(b->d_reps, b->d_items, b->nitems, b->d_queue, b->shape, true,
                                   b->kelbg, b->grid, b->stream);
    else {
        // two persistent launches, each with its own queue (d_queue[0..1], [2..3]);
        // they touch disjoint replicas, so the second may run beside the first
        const int nfull = b->nitems - b->nfine;
        const bool both = nfull > 0 && b->nfine > 0;
        if (both) {
            (void)hipEventRecord(b->fork, b->stream);
            (void)hipStreamWaitEvent(b->stream2, b->fork, 0);
        }
        if (nfull > 0)
            launch_pair_f64_batch(b->d_reps, b->d_items, nfull, b->d_queue, true, b->kelbg, b->grid,
                                  0, b->stream);
        if (b->nfine > 0)
            launch_pair_f64_batch(b->d_reps, b->d_items + nfull, b->nfine, b->d_queue + 2, true,
                                  b->kelbg, b->grid_fine, kFineN3Shape, both ? b->stream2 : b->stream);
        if (both) {
            (void)hipEventRecord(b->join, b->stream2);
            (void)hipStreamWaitEvent(b->stream, b->join, 0);
        }
    }
}

// One MD step of every replica, in the order of tcp_md_steps (main.f90:87-153).
void launch_step(tcp_batch* b, int thermo, double dt, double tau, double fric, uint64_t seed,
                 double* trace)
{
    launch_vv1_batch(b->d_reps, b->nrep, b->maxblk, thermo, dt, fric, b->stream);
    // Mixed batches run SPME on the second stream beside the pair launch
    // (forked after the kick-drift, joined before combine, which adds its
    // forces): its many short kernels fill what the persistent pair blocks
    // leave idle (64-replica batch 1.74 -> 1.71 ms per step).  The two touch
    // disjoint buffers, so every sum is unchanged.  The k-sum stays in line:
    // beside the pair launch its force kernel's blocks held CUs the pair
    // blocks then waited for (1.74 -> 1.87 ms).  FP64 batches keep both in
    // line: their second stream carries the 4 x 4 launch.
    const bool side = b->recip == TCP_RECIP_SPME && mixed(b);
    hipStream_t rs = side ? b->stream2 : b->stream;
    if (side) {
        (void)hipEventRecord(b->fork, b->stream);
        (void)hipStreamWaitEvent(b->stream2, b->fork, 0);
    }
    if (b->recip == TCP_RECIP_EWALD) launch_ewald_batch(b->d_reps, b->nrep, b->kmax, b->maxn, rs);
    if (b->recip == TCP_RECIP_SPME) launch_spme_batch(b->d_reps, b->spme, rs);
    if (side) (void)hipEventRecord(b->join, b->stream2);
    launch_pair(b);
    if (side) (void)hipStreamWaitEvent(b->stream, b->join, 0);
    launch_combine_batch(b->d_reps, b->nrep, b->maxblk, b->stream);
    launch_reduce_parts_batch(b->d_reps, b->nrep, 4, SC_FSUM, b->stream);
    launch_vv2_batch(b->d_reps, b->nrep, b->maxblk, thermo, dt, tau, fric, seed, b->d_counter,
                     b->stream);
    if (trace) launch_trace_batch(b->d_reps, b->nrep, trace, b->d_counter, b->stream);
    launch_step_advance(b->d_counter, b->stream);
}

int capture_step(tcp_batch* b, const tcp_batch::Key& k)
{
    if (b->exec && b->key == k) return 0;
    if (b->exec) {
        BHIP(b, hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    hipGraph_t g = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    launch_step(b, k.thermo, k.dt, k.tau, k.fric, k.seed, k.trace);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (le != hipSuccess) return bfail(b, TCP_ERR_HIP, std::string("launch in capture: ") + hipGetErrorString(le));
    if (ce != hipSuccess) return bfail(b, TCP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        b->exec = nullptr;
        return bfail(b, TCP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    b->key = k;
    return 0;
}

int check_rep(tcp_batch* b, int rep)
{
    if (rep < 0 || rep >= b->nrep) return bfail(b, TCP_ERR_ARG, "replica index out of range");
    return 0;
}

}  // namespace

extern "C" {

const char* tcp_batch_last_error(const tcp_batch* b)
{
    return b ? b->err.c_str() : g_batch_create_error.c_str();
}

int tcp_batch_destroy(tcp_batch* b)
{
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->stream2) (void)hipStreamSynchronize(b->stream2);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    for (auto e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (void* a : b->allocs) (void)hipFree(a);
    spme_batch_release(&b->spme);
    for (tcp_ctx* c : b->reps) tcp_destroy(c);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->stream2) (void)hipStreamDestroy(b->stream2);
    delete b;
    return 0;
}

int tcp_batch_create(tcp_batch** out, int nrep, const tcp_params* p, int precision)
{
    if (!out || !p || nrep < 1) return bfail(nullptr, TCP_ERR_ARG, "need an output pointer and nrep >= 1 parameter sets");
    if (nrep > 65535) return bfail(nullptr, TCP_ERR_ARG, "at most 65535 replicas (one grid row each)");
    *out = nullptr;
    for (int r = 1; r < nrep; ++r)
        if (p[r].recip != p[0].recip || p[r].flags != p[0].flags ||
            (p[0].recip == TCP_RECIP_EWALD && p[r].kmax != p[0].kmax) ||
            (p[0].recip == TCP_RECIP_SPME && p[r].order != p[0].order))
            return bfail(nullptr, TCP_ERR_ARG, "replicas must share recip, flags, kmax (k-sum) and order (SPME)");
    // batched SPME takes brick-binned grids only: the unbinned spreading adds
    // with FP64 atomics, whose order would break run-to-run repeatability
    for (int r = 0; r < nrep && p[0].recip == TCP_RECIP_SPME; ++r)
        for (int d = 0; d < 3; ++d)
            if (p[r].K[d] % kBrick != 0)
                return bfail(nullptr, TCP_ERR_UNSUPPORTED,
                             "replica " + std::to_string(r) + ": batches run SPME on grids with every K[d] a multiple of " +
                                 std::to_string(kBrick));
    for (int r = 0; r < nrep; ++r) {
        std::string msg;
        const int rc = validate_params(&p[r], precision, &msg);
        if (rc) return bfail(nullptr, rc, "replica " + std::to_string(r) + ": " + msg);
    }
    tcp_batch* b = new tcp_batch();
    b->nrep = nrep;
    b->precision = precision;
    b->recip = p[0].recip;
    b->kmax = p[0].kmax;
    b->kelbg = (p[0].flags & TCP_KELBG) != 0;
    auto bail = [&](int rc) {
        g_batch_create_error = b->err;
        tcp_batch_destroy(b);
        return rc;
    };
    for (int r = 0; r < nrep; ++r) {
        tcp_ctx* c = nullptr;
        const int rc = tcp_create(&c, &p[r], precision);
        if (rc) return bail(bfail(b, rc, "replica " + std::to_string(r) + ": " + tcp_last_error(nullptr)));
        b->reps.push_back(c);
    }
    if (hipGetDevice(&b->device) != hipSuccess ||
        hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess)
        return bail(bfail(b, TCP_ERR_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess)
        return bail(bfail(b, TCP_ERR_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
        return bail(bfail(b, TCP_ERR_HIP, "hipEventCreate failed"));
    for (auto& e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(bfail(b, TCP_ERR_HIP, "hipEventCreate failed"));
    int rc = 0;
    if ((rc = balloc(b, &b->d_reps, (size_t)nrep))) return bail(rc);
    if ((rc = balloc(b, &b->d_queue, (size_t)4))) return bail(rc);  // two persistent launches
    if ((rc = balloc(b, &b->d_counter, (size_t)2))) return bail(rc);
    if ((rc = balloc(b, &b->d_rdf_scratch, (size_t)3 * kHistStride))) return bail(rc);
    if (hipMemset(b->d_queue, 0, 4 * sizeof(int)) != hipSuccess ||
        hipMemset(b->d_counter, 0, 2 * sizeof(uint64_t)) != hipSuccess)
        return bail(bfail(b, TCP_ERR_HIP, "hipMemset failed"));
    b->h_reps.resize(nrep);
    b->pf.resize(nrep);
    b->pu.resize(nrep);
    for (int r = 0; r < nrep; ++r) {
        tcp_ctx* c = b->reps[r];
        int slots = 0;
        for (int sh = 0; sh < (mixed(b) ? kNumN3Shapes : 1); ++sh) slots = std::max(slots, slots_of(b, c, sh));
        if ((rc = balloc(b, &b->pf[r], (size_t)slots * 3 * c->N))) return bail(rc);
        if ((rc = balloc(b, &b->pu[r], (size_t)slots * c->N))) return bail(rc);
        b->maxblk = std::max(b->maxblk, md_blocks(c->N));
        b->maxn = std::max(b->maxn, c->N);
        BatchRep& h = b->h_reps[r];
        h = BatchRep{};
        h.m = md_args(c);
        h.pa = pair_args(c);
        h.pa.pi = PosSoA{c->x, c->y, c->z};
        h.pa.pj = h.pa.pi;
        h.pa.ui = c->upos;
        h.pa.uj = c->upos;
        h.pa.n_i = h.pa.n_j = c->N;
        h.pa.ne_i = h.pa.ne_j = c->Ne;
        h.pa.jchunk = c->jchunk;
        h.pa.rdf_weight = 1;
        h.pa.pf = b->pf[r];
        h.pa.pu = b->pu[r];
        h.ra = recip_args(c);
        h.sk = c->sk;
        h.part = c->part;
        h.sc = c->sc;
        h.recip = c->p.recip;
        h.kmax = c->p.kmax;
        h.alpha = c->alpha;
        h.ur_scale = 1.0;
        h.Uo = c->Uo;
        h.eps = c->p.eps;
        const double V = c->p.box[0] * c->p.box[1] * c->p.box[2];
        h.um_scale = c->p.eps / (2.0 * kPi * V);
    }
    if (b->recip == TCP_RECIP_SPME) {
        std::vector<RecipArgs> ra(nrep);
        for (int r = 0; r < nrep; ++r) ra[r] = b->h_reps[r].ra;
        if ((rc = spme_batch_plan(ra.data(), nrep, &b->spme)))
            return bail(bfail(b, rc, "SPME pass tables: hipMalloc failed"));
    }
    if ((rc = set_shape(b, 0))) return bail(rc);
    *out = b;
    return 0;
}

int tcp_batch_set_state(tcp_batch* b, int rep, const double* x, const double* v, const double* f)
{
    if (!b) return TCP_ERR_ARG;
    BTRY(check_rep(b, rep));
    const int rc = tcp_set_state(b->reps[rep], x, v, f);
    return rc ? bfail(b, rc, tcp_last_error(b->reps[rep])) : 0;
}

int tcp_batch_get_state(tcp_batch* b, int rep, double* x, double* v, double* f)
{
    if (!b) return TCP_ERR_ARG;
    BTRY(check_rep(b, rep));
    const int rc = tcp_get_state(b->reps[rep], x, v, f);
    return rc ? bfail(b, rc, tcp_last_error(b->reps[rep])) : 0;
}

int tcp_batch_get_histograms(tcp_batch* b, int rep, double* rdf, double* vdf, int reset)
{
    if (!b) return TCP_ERR_ARG;
    BTRY(check_rep(b, rep));
    const int rc = tcp_get_histograms(b->reps[rep], rdf, vdf, reset);
    return rc ? bfail(b, rc, tcp_last_error(b->reps[rep])) : 0;
}

int tcp_batch_set_shape(tcp_batch* b, int shape)
{
    if (!b) return TCP_ERR_ARG;
    if (shape < 0 || shape >= kNumN3Shapes || (!mixed(b) && shape != 0))
        return bfail(b, TCP_ERR_ARG, "unknown block shape");
    return set_shape(b, shape);
}

int tcp_batch_replica_shape(tcp_batch* b, int rep, int* shape)
{
    if (!b || !shape) return TCP_ERR_ARG;
    BTRY(check_rep(b, rep));
    *shape = rep_shape(b, b->reps[rep], b->shape);
    return 0;
}

int tcp_batch_autotune(tcp_batch* b, int reps, int* best, double* ms)
{
    if (!b || reps < 1) return TCP_ERR_ARG;
    // time on the current positions with the RDF counts sent to a scratch histogram
    std::vector<BatchRep> saved = b->h_reps;
    for (auto& h : b->h_reps) h.pa.rdf = b->d_rdf_scratch;
    const int nshape = mixed(b) ? kNumN3Shapes : 1;
    int arg = 0;
    double bestms = 1e300;
    for (int sh = 0; sh < nshape; ++sh) {
        BTRY(set_shape(b, sh));
        launch_pair(b);  // warm-up
        BHIP(b, hipEventRecord(b->ev[0], b->stream));
        for (int k = 0; k < reps; ++k) launch_pair(b);
        BHIP(b, hipEventRecord(b->ev[1], b->stream));
        BHIP(b, hipGetLastError());
        BHIP(b, hipEventSynchronize(b->ev[1]));
        float t = 0;
        BHIP(b, hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
        const double per = t / reps;
        if (ms) ms[sh] = per;
        if (per < bestms) {
            bestms = per;
            arg = sh;
        }
    }
    b->h_reps = saved;
    BTRY(set_shape(b, arg));
    b->pair_ms = bestms;
    if (best) *best = arg;
    return 0;
}

int tcp_batch_md_steps(tcp_batch* b, int nsteps, int thermo, double dt, const double* Te,
                       const double* Ti, double tau, double fric, uint64_t seed, double* trace)
{
    if (!b || nsteps < 0 || thermo < TCP_NOTEMP || thermo > TCP_STOCH) return TCP_ERR_ARG;
    if ((thermo != TCP_NOTEMP) && (!Te || !Ti))
        return bfail(b, TCP_ERR_ARG, "thermostats need per-replica Te and Ti");
    if (nsteps == 0) return 0;
    for (int r = 0; r < b->nrep; ++r) {
        b->h_reps[r].Te = Te ? Te[r] : 0.0;
        b->h_reps[r].Ti = Ti ? Ti[r] : 0.0;
    }
    BHIP(b, hipMemcpyAsync(b->d_reps, b->h_reps.data(), b->nrep * sizeof(BatchRep),
                           hipMemcpyHostToDevice, b->stream));
    if (trace && (size_t)nsteps * b->nrep > b->trace_cap) {
        BTRY(bgrow(b, &b->d_trace, (size_t)6 * nsteps * b->nrep));
        b->trace_cap = (size_t)nsteps * b->nrep;
    }
    const uint64_t ctr[2] = {b->counter, 0};
    BHIP(b, hipMemcpyAsync(b->d_counter, ctr, sizeof(ctr), hipMemcpyHostToDevice, b->stream));
    tcp_batch::Key k;
    k.thermo = thermo;
    k.dt = dt;
    k.tau = tau;
    k.fric = fric;
    k.seed = seed;
    k.trace = trace ? b->d_trace : nullptr;
    k.shape = b->shape;
    k.grid = b->grid;
    k.nitems = b->nitems;
    k.nfine = b->nfine;
    k.grid_fine = b->grid_fine;
    BTRY(capture_step(b, k));
    BHIP(b, hipEventRecord(b->ev[2], b->stream));
    for (int s = 0; s < nsteps; ++s) BHIP(b, hipGraphLaunch(b->exec, b->stream));
    BHIP(b, hipEventRecord(b->ev[3], b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    float ms = 0;
    BHIP(b, hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
    b->steps_ms = ms;
    b->counter += (uint64_t)nsteps;
    if (trace)
        BHIP(b, hipMemcpy(trace, b->d_trace, (size_t)6 * nsteps * b->nrep * sizeof(double),
                          hipMemcpyDeviceToHost));
    return 0;
}

int tcp_batch_last_timings(tcp_batch* b, double* pair_ms, double* steps_ms)
{
    if (!b) return TCP_ERR_ARG;
    if (pair_ms) *pair_ms = b->pair_ms;
    if (steps_ms) *steps_ms = b->steps_ms;
    return 0;
}

}  // extern "C"

// C ABI of libtcp_gpu (include/tcp_gpu.h): context, device memory and the
// orchestration of the kernels in kernels_pair.hip, kernels_spme.hip and
// kernels_md.hip.  Device layout: positions, velocities and forces as FP64
// structure of arrays (electrons first), a fixed-point position mirror (FixPos)
// for the mixed pair kernel, per-(chunk, particle) FP64 pair partials, the
// K1*K2*K3 SPME grids, and a small block of device scalars so that a whole MD
// step runs without a host round trip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "tcp_ctx.h"

using namespace tcp;

namespace tcp {

CopyPool::CopyPool(int workers) : parts_(workers + 1), pid_(getpid())
{
    for (int w = 1; w <= workers; ++w) th_.emplace_back(&CopyPool::work, this, w);
}

CopyPool& CopyPool::shared()
{
    static CopyPool* pool = new CopyPool(3);
    return *pool;
}

// part id of every job: bytes [bytes id / parts, bytes (id + 1) / parts)
void CopyPool::work(int id)
{
    unsigned seen = 0;
    for (;;) {
        char* d;
        const char* src;
        size_t n;
        {
            std::unique_lock<std::mutex> l(mu_);
            go_.wait(l, [&] { return gen_ != seen; });
            seen = gen_;
            d = dst_;
            src = src_;
            n = bytes_;
        }
        const size_t lo = n * id / parts_, hi = n * (id + 1) / parts_;
        std::memcpy(d + lo, src + lo, hi - lo);
        std::lock_guard<std::mutex> l(mu_);
        if (--left_ == 0) done_.notify_one();
    }
}

void CopyPool::copy(void* dst, const void* src, size_t bytes)
{
    // a forked child inherits the pool object but not its threads (and maybe a
    // held lock): it copies on the calling thread alone
    if (getpid() != pid_) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::lock_guard<std::mutex> job(job_);
    {
        std::lock_guard<std::mutex> l(mu_);
        dst_ = (char*)dst;
        src_ = (const char*)src;
        bytes_ = bytes;
        left_ = parts_ - 1;
        ++gen_;
    }
    go_.notify_all();
    std::memcpy(dst, src, bytes / parts_);
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [&] { return left_ == 0; });
}

}  // namespace tcp

namespace {

constexpr double kPi = 3.1415926535897931;      // data.f90:23
constexpr double kSqrtPi = 1.7724538509055159;  // data.f90:24

thread_local std::string g_create_error;

}  // namespace

namespace {

int fail(tcp_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCK(ctx, call)                                                                      \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, TCP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCK(ctx) HIPCK(ctx, hipGetLastError())

template <typename T>
int dalloc(tcp_ctx* c, T** p, size_t count)
{
    void* q = nullptr;
    HIPCK(c, hipMalloc(&q, count * sizeof(T) + 64));
    c->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

#define TRY(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_) return rc_;       \
    } while (0)

// Grows a buffer of the context to at least count elements: the old one is
// freed (after the stream has drained, since queued kernels may read it) and
// its contents are not kept.  *cap counts elements.
template <typename T>
int dgrow(tcp_ctx* c, T** p, size_t* cap, size_t count)
{
    if (*p && count <= *cap) return 0;
    if (*p) {
        HIPCK(c, hipStreamSynchronize(c->stream));
        auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)*p);
        if (it != c->allocs.end()) c->allocs.erase(it);
        HIPCK(c, hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    TRY(dalloc(c, p, count));
    *cap = count;
    return 0;
}

bool pow2(int k) { return k >= 2 && (k & (k - 1)) == 0; }

}  // namespace

namespace tcp {

MdArgs md_args(tcp_ctx* c)
{
    MdArgs m{};
    m.n = c->N;
    m.ne = c->Ne;
    m.x = c->x; m.y = c->y; m.z = c->z;
    m.vx = c->vx; m.vy = c->vy; m.vz = c->vz;
    m.fx = c->fx; m.fy = c->fy; m.fz = c->fz;
    m.upos = c->upos;
    for (int d = 0; d < 3; ++d) m.box[d] = c->p.box[d];
    m.xm[0] = c->p.xm[0];
    m.xm[1] = c->p.xm[1];
    m.dv[0] = c->p.dv_e;
    m.dv[1] = c->p.dv_i;
    m.vdf = c->vdf;
    return m;
}

RecipArgs recip_args(tcp_ctx* c)
{
    RecipArgs a{};
    a.n = c->N;
    a.ne = c->Ne;
    a.x = c->x; a.y = c->y; a.z = c->z;
    for (int d = 0; d < 3; ++d) {
        a.K[d] = c->p.K[d];
        a.box[d] = c->p.box[d];
        a.tw[d] = c->tw[d];
    }
    a.order = c->p.order;
    a.q[0] = c->p.q[0];
    a.q[1] = c->p.q[1];
    a.eps = c->p.eps;
    a.Q = c->Q; a.A = c->A; a.C = c->C; a.gb = c->gb;
    a.epart = c->epart;
    a.nepart = c->nepart;
    a.frx = c->frx; a.fry = c->fry; a.frz = c->frz;
    a.skpart = c->skpart;
    for (int d = 0; d < 3; ++d) a.nbrick[d] = c->nbrick[d];
    a.brick_of = c->brick_of;
    a.brick_count = c->brick_count;
    a.brick_start = c->brick_start;
    a.brick_cursor = c->brick_cursor;
    a.sorted = c->sorted;
    a.tiles = c->tiles;
    return a;
}

void choose_chunks(int n_i, int n_j, int* nchunk, int* jchunk)
{
    const int iblocks = (n_i + kPairBlock - 1) / kPairBlock;
    int nc = (2048 + iblocks - 1) / iblocks;             // aim at >= 2048 blocks
    const int max_by_size = (n_j + 127) / 128;           // at least 128 j per chunk
    if (nc > max_by_size) nc = max_by_size;
    if (nc < 1) nc = 1;
    if (nc > 64) nc = 64;
    int jc = (n_j + nc - 1) / nc;
    nc = (n_j + jc - 1) / jc;
    *nchunk = nc;
    *jchunk = jc;
}

PairArgs pair_args(tcp_ctx* c)
{
    PairArgs a{};
    for (int d = 0; d < 3; ++d) {
        a.box[d] = c->p.box[d];
        a.inv_box[d] = 1.0 / c->p.box[d];
    }
    a.rc2 = c->rc2;
    a.alpha = c->alpha;
    a.rs = c->p.rs;
    a.eps = c->p.eps;
    a.q[0] = c->p.q[0];
    a.q[1] = c->p.q[1];
    const double dx[3] = {c->p.dx_ee, c->p.dx_ei, c->p.dx_ii};
    for (int t = 0; t < 3; ++t) {
        a.dx[t] = dx[t];
        a.inv_dx[t] = 1.0 / dx[t];
    }
    a.rdf = c->rdf;
    a.flags = c->p.flags;
    return a;
}

}  // namespace tcp

namespace {

void launch_pair(tcp_ctx* c, const PairArgs& a, int nchunk)
{
    if (c->precision == TCP_MIXED) launch_pair_mixed(a, nchunk, c->stream);
    else launch_pair_f64(a, nchunk, c->stream);
}

// Real-space all-pair forces of the whole system into f, block partials into part.
// Reciprocal forces into frx/fry/frz and Um into sc[SC_UM] on stream s (null:
// the context's stream); with qp the per-particle charges replace the species
// charges, pot receives dUm/dq.
int recip(tcp_ctx* c, const double* qp = nullptr, double* pot = nullptr, hipStream_t s = nullptr);

// with_recip: the reciprocal part (recip) runs on stream2 beside the force sum
// and its forces join after it (SPME_SUM), with the same result as frx etc.
int direct_all(tcp_ctx* c, int rdf_weight, const double* frx, const double* fry, const double* frz,
               int ev_slot, bool with_recip = false)
{
    PairArgs a = pair_args(c);
    a.pi = PosSoA{c->x, c->y, c->z};
    a.pj = a.pi;
    a.ui = c->upos;
    a.uj = c->upos;
    a.n_i = a.n_j = c->N;
    a.ne_i = a.ne_j = c->Ne;
    a.jchunk = c->jchunk;
    a.rdf_weight = rdf_weight;
    a.pf = c->pf;
    a.pu = c->pu;
    if (ev_slot >= 0) HIPCK(c, hipEventRecord(c->ev_pair[0][ev_slot], c->stream));
    if (c->precision == TCP_MIXED)
        launch_pair_mixed_n3(a, c->blkmap, c->n3_blocks, c->n3_queue, c->n3_grid, c->n3_shape,
                             c->stream);
    else
        launch_pair_f64_n3(a, c->blkmap, c->n3_blocks, c->n3_queue, c->n3_grid, c->n3_shape,
                           c->stream);
    LAUNCHCK(c);
    if (ev_slot >= 0) HIPCK(c, hipEventRecord(c->ev_pair[1][ev_slot], c->stream));
    if (with_recip) {  // positions are final: the reciprocal part runs beside the sum
        HIPCK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCK(c, hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
        TRY(recip(c, nullptr, nullptr, c->stream2));
        HIPCK(c, hipEventRecord(c->ev_join, c->stream2));
    }
    // the Newton kernels give each unordered pair's energy once: Ur = sum u
    c->ur_scale = 1.0;
    launch_combine(md_args(c), c->pf, c->pu, c->n3_rows, n3_shape(c->n3_shape).tiles * 64,
                   with_recip ? nullptr : frx, with_recip ? nullptr : fry,
                   with_recip ? nullptr : frz, c->part, false, c->stream);
    LAUNCHCK(c);
    if (with_recip) {
        HIPCK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
        launch_add_recip(md_args(c), c->frx, c->fry, c->frz, c->part, c->stream);
        LAUNCHCK(c);
    }
    launch_reduce_parts(c->part, md_blocks(c->N), 4, c->sc + SC_FSUM, c->stream);
    LAUNCHCK(c);
    return 0;
}

int recip(tcp_ctx* c, const double* qp, double* pot, hipStream_t s)
{
    if (!s) s = c->stream;
    RecipArgs a = recip_args(c);
    a.qp = qp;
    a.pot = pot;
    const double V = c->p.box[0] * c->p.box[1] * c->p.box[2];
    if (c->p.recip == TCP_RECIP_SPME) {
        launch_spme_spread(a, s);
        LAUNCHCK(c);
        launch_spme_convolve(a, s);
        LAUNCHCK(c);
        launch_spme_gather(a, s);
        LAUNCHCK(c);
        launch_um(c->epart, c->nepart, c->p.eps / (2.0 * kPi * V), c->sc + SC_UM, s);
    } else {
        launch_ewald_kspace(a, c->p.kmax, c->alpha, c->sk, s);
        LAUNCHCK(c);
        launch_um(c->epart, 1, c->p.eps / (2.0 * kPi * V), c->sc + SC_UM, s);
    }
    LAUNCHCK(c);
    return 0;
}

int sync(tcp_ctx* c)
{
    HIPCK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int get_sc(tcp_ctx* c, int slot, int count, double* out)
{
    HIPCK(c, hipMemcpyAsync(out, c->sc + slot, count * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    return sync(c);
}

int vv2(tcp_ctx* c, int thermo, double dt, double Te, double Ti, double tau, double fric,
        uint64_t seed, const double* fsum)
{
    MdArgs m = md_args(c);
    const int nb = md_blocks(c->N);
    launch_vv2_kick(m, thermo, dt, Te, Ti, fric, seed, c->counter++, fsum, c->part, c->stream);
    LAUNCHCK(c);
    if (thermo == TCP_BEREND || thermo == TCP_ISOKIN) {
        launch_reduce_parts(c->part, nb, 2, c->sc + SC_MV2, c->stream);
        launch_thermo_lambda(c->sc + SC_MV2, thermo, c->Ne, c->N - c->Ne, Te, Ti, dt, tau,
                             c->sc + SC_LAM, c->stream);
        launch_vv2_scale(m, c->sc + SC_LAM, c->part, c->stream);
        LAUNCHCK(c);
    }
    launch_reduce_parts(c->part, nb, 2, c->sc + SC_MV2, c->stream);
    LAUNCHCK(c);
    return 0;
}

int ensure_ext(tcp_ctx* c, int ng)
{
    if (ng <= c->g_cap) return 0;
    double** arrs[6] = {&c->gx, &c->gy, &c->gz, &c->gfx, &c->gfy, &c->gfz};
    size_t cap = c->g_cap;
    for (auto a : arrs) {
        cap = c->g_cap;
        TRY(dgrow(c, a, &cap, (size_t)ng));
    }
    cap = c->g_cap;
    TRY(dgrow(c, &c->gupos, &cap, (size_t)ng));
    c->g_cap = ng;
    return 0;
}

// Host state in AoS (x0 y0 z0 x1 ...) moves through one pinned staging buffer
// and one device buffer, each large enough for the arrays of one call: a host
// copy into pinned memory, one DMA, and the transposes on the device (and the
// reverse for downloads).
int ensure_stage(tcp_ctx* c, size_t count)
{
    if (count <= c->stage_cap) return 0;
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->d_stage) (void)hipFree(c->d_stage);
    c->h_stage = nullptr;
    c->d_stage = nullptr;
    c->stage_cap = 0;
    if (hipHostMalloc((void**)&c->h_stage, count * sizeof(double)) != hipSuccess) {
        c->h_stage = nullptr;
        return fail(c, TCP_ERR_HIP, "hipHostMalloc of the host staging buffer failed");
    }
    if (hipMalloc((void**)&c->d_stage, count * sizeof(double)) != hipSuccess) {
        c->d_stage = nullptr;
        return fail(c, TCP_ERR_HIP, "hipMalloc of the device staging buffer failed");
    }
    c->stage_cap = count;
    return 0;
}

// One AoS host array and its device SoA triple.
struct Aos3 {
    double* host;
    double *x, *y, *z;
};

// The k arrays of one call, concatenated, move in at most kStageChunks pieces
// of at least 128 Ki doubles (1 MiB): the host copy of one piece overlaps the
// DMA of the previous one.
size_t stage_chunk(size_t total)
{
    const size_t by_count = (total + kStageChunks - 1) / kStageChunks;
    return by_count > ((size_t)1 << 17) ? by_count : ((size_t)1 << 17);
}

// doubles [off, off + len) of the concatenated host arrays, into (in) or out
// of the staging buffer; runs of 256 Ki bytes or more go through the pool
void host_piece(const Aos3* t, size_t m, size_t off, size_t len, double* stage, bool in)
{
    while (len > 0) {
        const size_t j = off / m, o = off % m, l = len < m - o ? len : m - o;
        void* d = in ? (void*)stage : (void*)(t[j].host + o);
        const void* src = in ? (const void*)(t[j].host + o) : (const void*)stage;
        if (l * sizeof(double) >= ((size_t)1 << 18)) CopyPool::shared().copy(d, src, l * sizeof(double));
        else std::memcpy(d, src, l * sizeof(double));
        stage += l;
        off += l;
        len -= l;
    }
}

int upload_aos(tcp_ctx* c, const Aos3* t, int k, int n)
{
    if (k == 0 || n == 0) return 0;
    const size_t m = 3 * (size_t)n, total = m * k, chunk = stage_chunk(total);
    TRY(ensure_stage(c, total));
    for (size_t off = 0; off < total; off += chunk) {
        const size_t len = total - off < chunk ? total - off : chunk;
        host_piece(t, m, off, len, c->h_stage + off, true);
        HIPCK(c, hipMemcpyAsync(c->d_stage + off, c->h_stage + off, len * sizeof(double),
                                hipMemcpyHostToDevice, c->stream));
    }
    for (int j = 0; j < k; ++j) launch_aos_to_soa(c->d_stage + m * j, t[j].x, t[j].y, t[j].z, n, c->stream);
    LAUNCHCK(c);
    return sync(c);
}

int download_aos(tcp_ctx* c, const Aos3* t, int k, int n)
{
    if (k == 0 || n == 0) return 0;
    const size_t m = 3 * (size_t)n, total = m * k, chunk = stage_chunk(total);
    TRY(ensure_stage(c, total));
    for (int j = 0; j < k; ++j) launch_soa_to_aos(t[j].x, t[j].y, t[j].z, c->d_stage + m * j, n, c->stream);
    LAUNCHCK(c);
    int e = 0;
    for (size_t off = 0; off < total; off += chunk, ++e) {
        const size_t len = total - off < chunk ? total - off : chunk;
        HIPCK(c, hipMemcpyAsync(c->h_stage + off, c->d_stage + off, len * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipEventRecord(c->ev_stage[e], c->stream));
    }
    e = 0;
    for (size_t off = 0; off < total; off += chunk, ++e) {
        const size_t len = total - off < chunk ? total - off : chunk;
        HIPCK(c, hipEventSynchronize(c->ev_stage[e]));
        host_piece(t, m, off, len, c->h_stage + off, false);
    }
    return 0;
}

int upload3(tcp_ctx* c, const double* aos, double* dx, double* dy, double* dz, int n)
{
    const Aos3 t{const_cast<double*>(aos), dx, dy, dz};
    return upload_aos(c, &t, 1, n);
}

int download3(tcp_ctx* c, double* aos, double* dx, double* dy, double* dz, int n)
{
    const Aos3 t{aos, dx, dy, dz};
    return download_aos(c, &t, 1, n);
}

int setup_recip(tcp_ctx* c)
{
    const int* K = c->p.K;
    const size_t G = (size_t)K[0] * K[1] * K[2];
    TRY(dalloc(c, &c->frx, (size_t)c->N));
    TRY(dalloc(c, &c->fry, (size_t)c->N));
    TRY(dalloc(c, &c->frz, (size_t)c->N));
    if (c->p.recip == TCP_RECIP_EWALD) {
        const int nk = 2 * c->p.kmax + 1;
        TRY(dalloc(c, &c->sk, (size_t)3 * nk * nk * nk));
        const int nchunk = (c->N + ewald_chunk(c->p.kmax) - 1) / ewald_chunk(c->p.kmax);
        TRY(dalloc(c, &c->skpart, (size_t)nchunk * ewald_half(c->p.kmax)));
        TRY(dalloc(c, &c->epart, (size_t)1));
        c->nepart = 1;
        return 0;
    }
    TRY(dalloc(c, &c->Q, G));
    TRY(dalloc(c, &c->C, G));
    TRY(dalloc(c, &c->gb, G));
    TRY(dalloc(c, &c->A, G));
    c->nepart = fft_epart_blocks(K);
    TRY(dalloc(c, &c->epart, (size_t)c->nepart));
    if (K[0] % kBrick == 0 && K[1] % kBrick == 0 && K[2] % kBrick == 0) {
        for (int d = 0; d < 3; ++d) c->nbrick[d] = K[d] / kBrick;
        const int nb = c->nbrick[0] * c->nbrick[1] * c->nbrick[2];
        const size_t E = kBrick + c->p.order - 1;
        TRY(dalloc(c, &c->brick_of, (size_t)c->N));
        TRY(dalloc(c, &c->sorted, (size_t)c->N));
        TRY(dalloc(c, &c->brick_count, (size_t)nb));
        TRY(dalloc(c, &c->brick_start, (size_t)nb + 1));
        TRY(dalloc(c, &c->brick_cursor, (size_t)nb));
        TRY(dalloc(c, &c->tiles, (size_t)nb * E * E * E));
    }
    // influence function PSI (force.f90:27-42, PSI(1,1,1) = 0) times the spline
    // modulus |b1 b2 b3|^2 (force.f90:834-839 for order 4), built on the host in FP64
    std::vector<double> bm[3];
    const int P = c->p.order;
    for (int d = 0; d < 3; ++d) {
        bm[d].resize(K[d]);
        // M_P(k+1), k = 0..P-2, by the cardinal-spline recursion at f = 0
        std::vector<double> W(P + 1, 0.0);
        W[0] = 0.0;  // M_2(0)
        W[1] = 1.0;  // M_2(1)
        for (int k = 3; k <= P; ++k)
            for (int j = k - 1; j >= 0; --j)
                W[j] = (j * W[j] + (k - j) * (j > 0 ? W[j - 1] : 0.0)) / (k - 1);
        for (int m = 0; m < K[d]; ++m) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k <= P - 2; ++k) {
                const double th = 2.0 * kPi * (double)m * k / K[d];
                re += W[k + 1] * std::cos(th);
                im += W[k + 1] * std::sin(th);
            }
            bm[d][m] = 1.0 / (re * re + im * im);
        }
    }
    std::vector<double> h(G);
    const double a2 = c->alpha * c->alpha, pi2 = kPi * kPi;
    for (int k = 0; k < K[2]; ++k) {
        const int iz = (k + K[2] / 2) % K[2] - K[2] / 2;
        for (int j = 0; j < K[1]; ++j) {
            const int iy = (j + K[1] / 2) % K[1] - K[1] / 2;
            for (int i = 0; i < K[0]; ++i) {
                const int ix = (i + K[0] / 2) % K[0] - K[0] / 2;
                const size_t g = ((size_t)k * K[1] + j) * K[0] + i;
                if (ix == 0 && iy == 0 && iz == 0) { h[g] = 0.0; continue; }
                const double m2 = (double)ix * ix / (c->p.box[0] * c->p.box[0]) +
                                  (double)iy * iy / (c->p.box[1] * c->p.box[1]) +
                                  (double)iz * iz / (c->p.box[2] * c->p.box[2]);
                h[g] = std::exp(-(m2 * pi2 / a2)) / m2 * bm[0][i] * bm[1][j] * bm[2][k];
            }
        }
    }
    HIPCK(c, hipMemcpy(c->gb, h.data(), G * sizeof(double), hipMemcpyHostToDevice));
    for (int d = 0; d < 3; ++d) {
        std::vector<double2> t(K[d] / 2);
        for (int m = 0; m < K[d] / 2; ++m) {
            // exp(+2 pi i m/K) from long double to round once
            const long double th = 2.0L * 3.14159265358979323846264338327950288L * m / K[d];
            t[m] = make_double2((double)cosl(th), (double)sinl(th));
        }
        TRY(dalloc(c, &c->tw[d], t.size()));
        HIPCK(c, hipMemcpy(c->tw[d], t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    return 0;
}

}  // namespace

namespace tcp {

int set_create_error(int code, const char* msg)
{
    g_create_error = msg;
    return code;
}

int validate_params(const tcp_params* p, int precision, std::string* msg)
{
    auto vfail = [](std::string* m, int code, const char* text) {
        *m = text;
        return code;
    };
    if (p->ne < 0 || p->ni < 0 || (int64_t)p->ne + p->ni < 2)
        return vfail(msg, TCP_ERR_ARG, "need at least two particles");
    // the block-row partials take N^2/16 bytes (275 GB at 2^21 particles)
    if ((int64_t)p->ne + p->ni > kMaxParticles)
        return vfail(msg, TCP_ERR_ARG, "at most 2^22 particles per context");
    if (!(p->xm[0] > 0) || !(p->xm[1] > 0))
        return vfail(msg, TCP_ERR_ARG, "masses must be positive");
    if (!(p->dv_e >= 0) || !(p->dv_i >= 0))
        return vfail(msg, TCP_ERR_ARG, "VDF bin widths must not be negative");
    if (p->flags & ~(TCP_KELBG | TCP_REG_EE))
        return vfail(msg, TCP_ERR_ARG, "unknown regularisation flags");
    if (precision != TCP_FP64 && precision != TCP_MIXED)
        return vfail(msg, TCP_ERR_ARG, "precision must be TCP_FP64 or TCP_MIXED");
    for (int d = 0; d < 3; ++d)
        if (!(p->box[d] > 0)) return vfail(msg, TCP_ERR_ARG, "box lengths must be positive");
    if (!(p->rc > 0) || !(p->rs > 0) || !(p->dx_ee > 0) || !(p->dx_ei > 0) || !(p->dx_ii > 0))
        return vfail(msg, TCP_ERR_ARG, "rc, rs and the RDF bin widths must be positive");
    const double rc_cut = std::fmin(std::fmin(p->box[0], p->box[1]), p->box[2]) / 2.0;
    if (p->rc > rc_cut) return vfail(msg, TCP_ERR_CUTOFF, " *** error at cut-off radius *** ");
    if (std::isnan(p->alpha) || std::isinf(p->alpha))
        return vfail(msg, TCP_ERR_ARG, "alpha must be finite (<= 0 selects 3.1/rc)");
    // the mixed pair kernels' erfc(x) = exp(-x^2) erfcx(x) reads erfcx from a table
    // of rows up to just past x = alpha r = kMixedErfcMax; every pair they evaluate
    // has r < rc
    if (precision == TCP_MIXED) {
        const double a = p->alpha > 0 ? p->alpha : 3.1 / p->rc;
        if (a * p->rc > kMixedErfcMax)
            return vfail(msg, TCP_ERR_UNSUPPORTED,
                         "mixed precision needs alpha*rc <= 4.5, the range of its erfcx table");
    }
    if (p->recip == TCP_RECIP_SPME) {
        for (int d = 0; d < 3; ++d)
            if (!pow2(p->K[d]) || p->K[d] > 1024 || p->K[d] < p->order)
                return vfail(msg, TCP_ERR_UNSUPPORTED,
                            "SPME grid sizes must be powers of two in [order, 1024]");
        if (p->order < 4 || p->order > 12 || (p->order & 1))
            return vfail(msg, TCP_ERR_UNSUPPORTED, "B-spline order must be 4, 6, 8, 10 or 12");
    } else if (p->recip == TCP_RECIP_EWALD) {
        if (p->kmax < 1 || p->kmax > 40) return vfail(msg, TCP_ERR_ARG, "kmax must be in [1, 40]");
    } else if (p->recip != TCP_RECIP_NONE) {
        return vfail(msg, TCP_ERR_ARG, "unknown reciprocal method");
    }
    return 0;
}

}  // namespace tcp

extern "C" {

const char* tcp_last_error(const tcp_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int tcp_create(tcp_ctx** out, const tcp_params* p, int precision)
{
    if (!out || !p) return fail(nullptr, TCP_ERR_ARG, "null argument");
    *out = nullptr;
    {
        std::string msg;
        const int rc = validate_params(p, precision, &msg);
        if (rc) return fail(nullptr, rc, msg);
    }
    tcp_ctx* c = new tcp_ctx();
    c->p = *p;
    c->precision = precision;
    c->N = p->ne + p->ni;
    c->Ne = p->ne;
    c->alpha = p->alpha > 0 ? p->alpha : 3.1 / p->rc;  // force.f90:23
    c->rc2 = p->rc * p->rc;
    c->Uo = -((double)p->ne * (p->q[0] * p->q[0])) - (double)p->ni * (p->q[1] * p->q[1]);
    c->Uo = c->Uo * c->alpha / kSqrtPi;  // force.f90:44-45
    auto bail = [&](int rc) {
        g_create_error = c->err;
        tcp_destroy(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess) return bail(fail(c, TCP_ERR_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(c, TCP_ERR_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
        return bail(fail(c, TCP_ERR_HIP, "hipEventCreate failed"));
    for (int e = 0; e < kEvPool; ++e)
        for (int k = 0; k < 2; ++k)
            if (hipEventCreate(&c->ev_pair[k][e]) != hipSuccess)
                return bail(fail(c, TCP_ERR_HIP, "hipEventCreate failed"));
    if (hipEventCreate(&c->ev_all[0]) != hipSuccess || hipEventCreate(&c->ev_all[1]) != hipSuccess)
        return bail(fail(c, TCP_ERR_HIP, "hipEventCreate failed"));
    for (auto& e : c->ev_stage)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bail(fail(c, TCP_ERR_HIP, "hipEventCreate failed"));
    c->events = true;
    const size_t n = c->N;
    int rc = 0;
    double** arrs[9] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->fx, &c->fy, &c->fz};
    for (auto a : arrs)
        if ((rc = dalloc(c, a, n))) return bail(rc);
    if ((rc = dalloc(c, &c->upos, n))) return bail(rc);
    choose_chunks(c->N, c->N, &c->nchunk, &c->jchunk);
    int slots = c->nchunk;
    {
        int cus = 1;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
            return bail(fail(c, TCP_ERR_HIP, "no multiprocessor count"));
        const bool kel = (p->flags & TCP_KELBG) != 0;
        // 4 x 4 tile blocks for small systems (n3_self_shape: a function of
        // N, the precision and the CU count only)
        const bool mixed = precision == TCP_MIXED;
        c->n3_shape = n3_self_shape(c->N, cus * (mixed ? pair_n3_occupancy(true, kel, 0)
                                                       : pair_f64_n3_occupancy(true, kel, 0)));
        c->n3_grid = cus * (mixed ? pair_n3_occupancy(true, kel, c->n3_shape)
                                  : pair_f64_n3_occupancy(true, kel, c->n3_shape));
        // upper triangle of block pairs, off-diagonal (full) blocks first
        c->n3_rows = n3_block_rows_shape(c->N, c->n3_shape);
        std::vector<int2> map;
        for (int r = 0; r < c->n3_rows; ++r)
            for (int q = r + 1; q < c->n3_rows; ++q) map.push_back(make_int2(r, q));
        for (int r = 0; r < c->n3_rows; ++r) map.push_back(make_int2(r, r));
        c->n3_blocks = (int)map.size();
        if ((rc = dalloc(c, &c->blkmap, map.size()))) return bail(rc);
        if (hipMemcpy(c->blkmap, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(c, TCP_ERR_HIP, "hipMemcpy of the block map failed"));
        if ((rc = dalloc(c, &c->n3_queue, (size_t)2))) return bail(rc);
        if (hipMemset(c->n3_queue, 0, 2 * sizeof(int)) != hipSuccess)
            return bail(fail(c, TCP_ERR_HIP, "hipMemset of the work queue failed"));
        if (c->n3_rows > slots) slots = c->n3_rows;
    }
    c->pf_cap = (size_t)slots * 3 * n;
    if ((rc = dalloc(c, &c->pf, c->pf_cap))) return bail(rc);
    if ((rc = dalloc(c, &c->pu, (size_t)slots * n))) return bail(rc);
    if ((rc = dgrow(c, &c->part, &c->part_cap, (size_t)4 * md_blocks(c->N)))) return bail(rc);
    if ((rc = dalloc(c, &c->sc, (size_t)SC_N))) return bail(rc);
    if ((rc = dalloc(c, &c->rdf, (size_t)3 * kHistStride))) return bail(rc);
    if ((rc = dalloc(c, &c->vdf, (size_t)2 * kHistStride))) return bail(rc);
    if (p->recip != TCP_RECIP_NONE && (rc = setup_recip(c))) return bail(rc);
    // zero everything the kernels accumulate into (fill kernels, no memset nodes)
    if (hipMemsetAsync(c->rdf, 0, 3 * kHistStride * sizeof(unsigned long long), c->stream) != hipSuccess ||
        hipMemsetAsync(c->vdf, 0, 2 * kHistStride * sizeof(unsigned long long), c->stream) != hipSuccess ||
        hipMemsetAsync(c->sc, 0, SC_N * sizeof(double), c->stream) != hipSuccess)
        return bail(fail(c, TCP_ERR_HIP, "hipMemsetAsync failed"));
    double* zero3[9] = {c->x, c->y, c->z, c->vx, c->vy, c->vz, c->fx, c->fy, c->fz};
    for (auto a : zero3) launch_fill(a, n, 0.0, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(c, TCP_ERR_HIP, "initialisation failed"));
    *out = c;
    return 0;
}

int tcp_destroy(tcp_ctx* c)
{
    if (!c) return 0;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* a : c->allocs) (void)hipFree(a);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->d_stage) (void)hipFree(c->d_stage);
    if (c->events) {
        for (int e = 0; e < kEvPool; ++e)
            for (int k = 0; k < 2; ++k) (void)hipEventDestroy(c->ev_pair[k][e]);
        (void)hipEventDestroy(c->ev_all[0]);
        (void)hipEventDestroy(c->ev_all[1]);
        for (auto e : c->ev_stage) (void)hipEventDestroy(e);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    delete c;
    return 0;
}

int tcp_ewald_self(tcp_ctx* c, double* alpha, double* rc2, double* Uo, double* psi)
{
    if (!c) return TCP_ERR_ARG;
    if (alpha) *alpha = c->alpha;
    if (rc2) *rc2 = c->rc2;
    if (Uo) *Uo = c->Uo;
    if (psi) {
        const int* K = c->p.K;
        const double pi2 = kPi * kPi, a2 = c->alpha * c->alpha;
        for (int k = 0; k < K[2]; ++k) {
            const int iz = (k + K[2] / 2) % K[2] - K[2] / 2;
            for (int j = 0; j < K[1]; ++j) {
                const int iy = (j + K[1] / 2) % K[1] - K[1] / 2;
                for (int i = 0; i < K[0]; ++i) {
                    const int ix = (i + K[0] / 2) % K[0] - K[0] / 2;
                    const size_t g = ((size_t)k * K[1] + j) * K[0] + i;
                    if (ix == 0 && iy == 0 && iz == 0) { psi[g] = 0.0; continue; }
                    const double m2 = (double)ix * (double)ix / (c->p.box[0] * c->p.box[0]) +
                                      (double)iy * (double)iy / (c->p.box[1] * c->p.box[1]) +
                                      (double)iz * (double)iz / (c->p.box[2] * c->p.box[2]);
                    psi[g] = std::exp(-(m2 * pi2 / a2)) / m2;
                }
            }
        }
    }
    return 0;
}

int tcp_set_state(tcp_ctx* c, const double* x, const double* v, const double* f)
{
    if (!c) return TCP_ERR_ARG;
    Aos3 t[3];
    int k = 0;
    if (x) t[k++] = Aos3{const_cast<double*>(x), c->x, c->y, c->z};
    if (v) t[k++] = Aos3{const_cast<double*>(v), c->vx, c->vy, c->vz};
    if (f) t[k++] = Aos3{const_cast<double*>(f), c->fx, c->fy, c->fz};
    TRY(upload_aos(c, t, k, c->N));
    if (x) {
        launch_make_upos(md_args(c), c->stream);
        LAUNCHCK(c);
        TRY(sync(c));
    }
    return 0;
}

int tcp_get_state(tcp_ctx* c, double* x, double* v, double* f)
{
    if (!c) return TCP_ERR_ARG;
    Aos3 t[3];
    int k = 0;
    if (x) t[k++] = Aos3{x, c->x, c->y, c->z};
    if (v) t[k++] = Aos3{v, c->vx, c->vy, c->vz};
    if (f) t[k++] = Aos3{f, c->fx, c->fy, c->fz};
    TRY(download_aos(c, t, k, c->N));
    return 0;
}

int tcp_force_direct(tcp_ctx* c, double* Ur)
{
    if (!c) return TCP_ERR_ARG;
    TRY(direct_all(c, 1, nullptr, nullptr, nullptr, 0));
    double u;
    TRY(get_sc(c, SC_USUM, 1, &u));
    float ms = 0;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev_pair[0][0], c->ev_pair[1][0]));
    c->pair_ms = ms;
    if (Ur) *Ur = c->ur_scale * u;
    return 0;
}

int tcp_force_direct_ext(tcp_ctx* c, const double* xg, int ne_g, int ni_g, double* gemit,
                         double* Ur)
{
    if (!c || !xg || ne_g < 0 || ni_g < 0) return TCP_ERR_ARG;
    const int ng = ne_g + ni_g;
    if (ng == 0) {
        if (Ur) *Ur = 0.0;
        return 0;
    }
    TRY(ensure_ext(c, ng));
    TRY(upload3(c, xg, c->gx, c->gy, c->gz, ng));
    MdArgs mg = md_args(c);
    mg.n = ng;
    mg.ne = ne_g;
    mg.x = c->gx; mg.y = c->gy; mg.z = c->gz;
    mg.fx = c->gfx; mg.fy = c->gfy; mg.fz = c->gfz;
    mg.upos = c->gupos;
    launch_make_upos(mg, c->stream);
    // own (i) against external (j): f_own += cross forces, RDF +2 per pair
    int nc, jc;
    choose_chunks(c->N, ng, &nc, &jc);
    const size_t need = (size_t)nc * 3 * c->N;
    if (need > c->pf_cap) {
        size_t ucap = c->pf_cap / 3;
        TRY(dgrow(c, &c->pu, &ucap, (size_t)nc * c->N));
        TRY(dgrow(c, &c->pf, &c->pf_cap, need));
    }
    PairArgs a = pair_args(c);
    a.pi = PosSoA{c->x, c->y, c->z};
    a.pj = PosSoA{c->gx, c->gy, c->gz};
    a.ui = c->upos;
    a.uj = c->gupos;
    a.n_i = c->N; a.ne_i = c->Ne;
    a.n_j = ng; a.ne_j = ne_g;
    a.jchunk = jc;
    a.rdf_weight = 2;
    a.pf = c->pf;
    a.pu = c->pu;
    launch_pair(c, a, nc);
    LAUNCHCK(c);
    launch_combine(md_args(c), c->pf, c->pu, nc, 0, nullptr, nullptr, nullptr, c->part, true, c->stream);
    launch_reduce_parts(c->part, md_blocks(c->N), 4, c->sc + SC_FSUM, c->stream);
    LAUNCHCK(c);
    double u;
    TRY(get_sc(c, SC_USUM, 1, &u));
    // reaction forces on the external block (gemit, force.f90:237-337)
    int nc2, jc2;
    choose_chunks(ng, c->N, &nc2, &jc2);
    const size_t need2 = (size_t)nc2 * 3 * ng;
    if (need2 > c->gpf_cap) {
        size_t ucap = c->gpf_cap / 3;
        TRY(dgrow(c, &c->gpu, &ucap, (size_t)nc2 * ng));
        TRY(dgrow(c, &c->gpf, &c->gpf_cap, need2));
    }
    PairArgs b = a;
    b.pi = a.pj; b.pj = a.pi;
    b.ui = a.uj; b.uj = a.ui;
    b.n_i = ng; b.ne_i = ne_g;
    b.n_j = c->N; b.ne_j = c->Ne;
    b.jchunk = jc2;
    b.rdf_weight = 0;
    b.pf = c->gpf;
    b.pu = c->gpu;
    launch_pair(c, b, nc2);
    LAUNCHCK(c);
    // the block partials of the external block's combine: 4 per block of ng
    TRY(dgrow(c, &c->part, &c->part_cap, (size_t)4 * md_blocks(ng)));
    launch_combine(mg, c->gpf, c->gpu, nc2, 0, nullptr, nullptr, nullptr, c->part, false, c->stream);
    LAUNCHCK(c);
    if (gemit) TRY(download3(c, gemit, c->gfx, c->gfy, c->gfz, ng));
    TRY(sync(c));
    if (Ur) *Ur = u;
    return 0;
}

int tcp_force_recip(tcp_ctx* c, double* Um)
{
    if (!c) return TCP_ERR_ARG;
    if (c->p.recip == TCP_RECIP_NONE) {
        if (Um) *Um = 0.0;
        return 0;
    }
    TRY(recip(c));
    launch_add_forces(md_args(c), c->frx, c->fry, c->frz, c->stream);
    LAUNCHCK(c);
    double um;
    TRY(get_sc(c, SC_UM, 1, &um));
    if (Um) *Um = um;
    return 0;
}

int tcp_remove_rigidmtn(tcp_ctx* c)
{
    if (!c) return TCP_ERR_ARG;
    MdArgs m = md_args(c);
    launch_force_sum(m, c->part, c->stream);
    launch_reduce_parts(c->part, md_blocks(c->N), 3, c->sc + SC_FSUM, c->stream);
    launch_sub_mean(m, c->sc + SC_FSUM, c->stream);
    LAUNCHCK(c);
    return 0;  // device state only: the next call that returns host data waits
}

int tcp_vverlet1(tcp_ctx* c, int thermo, double dt, double fric)
{
    if (!c || thermo < TCP_NOTEMP || thermo > TCP_STOCH) return TCP_ERR_ARG;
    launch_vv1(md_args(c), thermo, dt, fric, c->stream);
    LAUNCHCK(c);
    return 0;  // device state only: the next call that returns host data waits
}

int tcp_vverlet2(tcp_ctx* c, int thermo, double dt, double Te, double Ti, double tau, double fric,
                 uint64_t seed, double* mv2el, double* mv2ion)
{
    if (!c || thermo < TCP_NOTEMP || thermo > TCP_STOCH) return TCP_ERR_ARG;
    TRY(vv2(c, thermo, dt, Te, Ti, tau, fric, seed, nullptr));
    double mv2[2];
    TRY(get_sc(c, SC_MV2, 2, mv2));
    if (mv2el) *mv2el = mv2[0];
    if (mv2ion) *mv2ion = mv2[1];
    return 0;
}

int tcp_vverlet2_kick(tcp_ctx* c, double dt, double* mv2el, double* mv2ion)
{
    if (!c) return TCP_ERR_ARG;
    // the BEREND/ISOKIN kick: half kick and the pre-rescale sums, no VDF
    launch_vv2_kick(md_args(c), TCP_BEREND, dt, 0.0, 0.0, 0.0, 0, 0, nullptr, c->part, c->stream);
    LAUNCHCK(c);
    launch_reduce_parts(c->part, md_blocks(c->N), 2, c->sc + SC_MV2, c->stream);
    LAUNCHCK(c);
    double mv2[2];
    TRY(get_sc(c, SC_MV2, 2, mv2));
    if (mv2el) *mv2el = mv2[0];
    if (mv2ion) *mv2ion = mv2[1];
    return 0;
}

int tcp_vverlet2_scale(tcp_ctx* c, double lam_e, double lam_i, double* mv2el, double* mv2ion)
{
    if (!c) return TCP_ERR_ARG;
    if (!std::isfinite(lam_e) || !std::isfinite(lam_i))
        return fail(c, TCP_ERR_ARG, "thermostat factors must be finite");
    const double lam[2] = {lam_e, lam_i};
    HIPCK(c, hipMemcpyAsync(c->sc + SC_LAM, lam, sizeof(lam), hipMemcpyHostToDevice, c->stream));
    launch_vv2_scale(md_args(c), c->sc + SC_LAM, c->part, c->stream);
    LAUNCHCK(c);
    launch_reduce_parts(c->part, md_blocks(c->N), 2, c->sc + SC_MV2, c->stream);
    LAUNCHCK(c);
    double mv2[2];
    TRY(get_sc(c, SC_MV2, 2, mv2));
    if (mv2el) *mv2el = mv2[0];
    if (mv2ion) *mv2ion = mv2[1];
    return 0;
}

int tcp_get_histograms(tcp_ctx* c, double* rdf, double* vdf, int reset)
{
    if (!c) return TCP_ERR_ARG;
    std::vector<unsigned long long> h(5 * kHistStride);
    HIPCK(c, hipMemcpyAsync(h.data(), c->rdf, 3 * kHistStride * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipMemcpyAsync(h.data() + 3 * kHistStride, c->vdf,
                            2 * kHistStride * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            c->stream));
    if (reset) {  // queued behind the copies on the same stream: one wait for both
        HIPCK(c, hipMemsetAsync(c->rdf, 0, 3 * kHistStride * sizeof(unsigned long long), c->stream));
        HIPCK(c, hipMemsetAsync(c->vdf, 0, 2 * kHistStride * sizeof(unsigned long long), c->stream));
    }
    TRY(sync(c));
    for (int t = 0; t < 3 && rdf; ++t)
        for (int b = 0; b < kNbin; ++b) rdf[t * kNbin + b] += (double)h[t * kHistStride + b];
    for (int t = 0; t < 2 && vdf; ++t)
        for (int b = 0; b < kNbin; ++b) vdf[t * kNbin + b] += (double)h[(3 + t) * kHistStride + b];
    return 0;
}

// systems from this size on run the reciprocal part beside the force sum
constexpr int kRecipBesideMinN = 65536;

int tcp_md_steps(tcp_ctx* c, int nsteps, int thermo, double dt, double Te, double Ti, double tau,
                 double fric, uint64_t seed, double* trace)
{
    if (!c || nsteps < 0 || thermo < TCP_NOTEMP || thermo > TCP_STOCH) return TCP_ERR_ARG;
    if (nsteps == 0) return 0;
    if (trace && nsteps > c->trace_cap) {
        size_t cap = (size_t)6 * c->trace_cap;
        TRY(dgrow(c, &c->trace, &cap, (size_t)6 * nsteps));
        c->trace_cap = nsteps;
    }
    const bool rec = c->p.recip != TCP_RECIP_NONE;
    MdArgs m = md_args(c);
    HIPCK(c, hipEventRecord(c->ev_all[0], c->stream));
    const int first_timed = nsteps > kEvPool ? nsteps - kEvPool : 0;
    for (int s = 0; s < nsteps; ++s) {
        launch_vv1(m, thermo, dt, fric, c->stream);  // main.f90:87-95
        LAUNCHCK(c);
        const int slot = s >= first_timed ? s - first_timed : -1;
        // main.f90:100 (the reciprocal part), 106, 136.  Large systems run the
        // reciprocal part beside the force sum; below kRecipBesideMinN the
        // sum is too short to hide it behind and the stream fork and join
        // would cost more than they save.
        const bool beside = rec && c->N >= kRecipBesideMinN;
        if (rec && !beside) TRY(recip(c));
        TRY(direct_all(c, 1, rec && !beside ? c->frx : nullptr, rec && !beside ? c->fry : nullptr,
                       rec && !beside ? c->frz : nullptr, slot, beside));
        TRY(vv2(c, thermo, dt, Te, Ti, tau, fric, seed, c->sc + SC_FSUM));  // main.f90:141-153
        if (trace) {
            launch_trace(c->trace + 6 * (size_t)s, c->sc + SC_USUM, c->ur_scale, c->Uo,
                         rec ? c->sc + SC_UM : nullptr, c->sc + SC_MV2, c->p.eps, c->stream);
            LAUNCHCK(c);
        }
    }
    HIPCK(c, hipEventRecord(c->ev_all[1], c->stream));
    TRY(sync(c));
    float ms = 0;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev_all[0], c->ev_all[1]));
    c->steps_ms = ms;
    double tot = 0;
    const int ntimed = nsteps - first_timed;
    for (int e = 0; e < ntimed; ++e) {
        HIPCK(c, hipEventElapsedTime(&ms, c->ev_pair[0][e], c->ev_pair[1][e]));
        tot += ms;
    }
    c->pair_ms = tot / ntimed;
    if (trace) {
        HIPCK(c, hipMemcpy(trace, c->trace, 6 * (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost));
    }
    return 0;
}

int tcp_set_lambda(tcp_ctx* c, double lambda)
{
    if (!c) return TCP_ERR_ARG;
    if (!(lambda > 0) || !std::isfinite(lambda)) return fail(c, TCP_ERR_ARG, "lambda must be positive");
    c->p.rs = lambda;
    return 0;
}

int tcp_energy_adjoint(tcp_ctx* c, const double* q, double* U, double* dUdx, double* dUdq,
                       double* dUdl, double* dFdl)
{
    if (!c) return TCP_ERR_ARG;
    if (!U || !dUdx || !dUdq || !dUdl) return fail(c, TCP_ERR_ARG, "null output");
    const int n = c->N;
    if (!c->adj_part) {
        TRY(dalloc(c, &c->adj_part, (size_t)kAdjQ * c->nchunk * n));
        TRY(dalloc(c, &c->adj_out, (size_t)kAdjQ * n));
        TRY(dalloc(c, &c->adj_q, (size_t)n));
        TRY(dalloc(c, &c->adj_pot, (size_t)n));
    }
    std::vector<double> qh(n);
    double q2 = 0.0;
    for (int i = 0; i < n; ++i) {
        qh[i] = q ? q[i] : c->p.q[i >= c->Ne];
        if (!std::isfinite(qh[i])) return fail(c, TCP_ERR_ARG, "charges must be finite");
        q2 += qh[i] * qh[i];
    }
    HIPCK(c, hipMemcpyAsync(c->adj_q, qh.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    AdjArgs a{};
    a.p = PosSoA{c->x, c->y, c->z};
    a.qp = c->adj_q;
    a.n = n;
    a.ne = c->Ne;
    a.jchunk = c->jchunk;
    a.flags = c->p.flags;
    for (int d = 0; d < 3; ++d) {
        a.box[d] = c->p.box[d];
        a.inv_box[d] = 1.0 / c->p.box[d];
    }
    a.rc2 = c->rc2;
    a.alpha = c->alpha;
    a.rs = c->p.rs;
    a.eps = c->p.eps;
    a.part = c->adj_part;
    launch_pair_adjoint(a, c->nchunk, c->stream);
    LAUNCHCK(c);
    launch_adjoint_reduce(c->adj_part, c->nchunk, n, c->adj_out, c->stream);
    LAUNCHCK(c);
    const bool rec = c->p.recip != TCP_RECIP_NONE;
    if (rec) TRY(recip(c, c->adj_q, c->adj_pot));
    std::vector<double> out((size_t)kAdjQ * n), fr(rec ? 4 * (size_t)n : 0);
    double um = 0.0;
    HIPCK(c, hipMemcpyAsync(out.data(), c->adj_out, out.size() * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream));
    if (rec) {
        const double* src[4] = {c->frx, c->fry, c->frz, c->adj_pot};
        for (int k = 0; k < 4; ++k)
            HIPCK(c, hipMemcpyAsync(fr.data() + (size_t)k * n, src[k], n * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
        HIPCK(c, hipMemcpyAsync(&um, c->sc + SC_UM, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    TRY(sync(c));
    // compensated sums of the per-particle energies in particle order
    auto nsum = [](double& s, double& e, double v) {
        const double t = s + v;
        e += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    };
    double us = 0.0, ue = 0.0, ls = 0.0, le = 0.0;
    const double self = 2.0 * c->alpha / kSqrtPi;
    for (int i = 0; i < n; ++i) {
        nsum(us, ue, out[3 * (size_t)n + i]);
        nsum(ls, le, out[5 * (size_t)n + i]);
        double dq = c->p.eps * (out[4 * (size_t)n + i] - self * qh[i]);
        if (rec) dq += fr[3 * (size_t)n + i];
        dUdq[i] = dq;
        for (int k = 0; k < 3; ++k) {
            double f = out[(size_t)k * n + i];
            if (rec) f += fr[(size_t)k * n + i];
            dUdx[3 * (size_t)i + k] = -f;
            if (dFdl) dFdl[3 * (size_t)i + k] = out[(size_t)(6 + k) * n + i];
        }
    }
    const double Uo = -q2 * c->alpha / kSqrtPi;
    *U = c->p.eps * (0.5 * (us + ue) + Uo) + um;
    *dUdl = c->p.eps * 0.5 * (ls + le);
    return 0;
}

int tcp_last_timings(tcp_ctx* c, double* pair_ms, double* steps_ms)
{
    if (!c) return TCP_ERR_ARG;
    if (pair_ms) *pair_ms = c->pair_ms;
    if (steps_ms) *steps_ms = c->steps_ms;
    return 0;
}

}  // extern "C"

// O(N) kernels of the step: velocity Verlet halves with the thermostats
// (vverlet.f90), the force sum SPME_SUM (force.f90:950-965), mean-force
// removal REMOVE_RIGIDMTN (force.f90:974-1007) and deterministic reductions.
//
// The per-particle updates keep the reference's operation order with FMA
// contraction off, so the FP64 integrator is bitwise the reference's for the
// same forces.  Reductions are two-level in a fixed order (per-block LDS tree,
// then one block over the block partials), so FP64 sums do not depend on
// scheduling.  Histogram counts use 32-bit LDS atomics per block and one
// 64-bit global atomic per non-empty bin.
#include "tcp_internal.h"

namespace tcp {

namespace {

// Block sum of W values per thread into part[blockIdx.x * W + c], fixed order.
template <int W, int S = W>  // S: the record length of part (W <= S)
__device__ inline void block_sum_store(double (&v)[W], double* part, int blk)
{
    __shared__ double red[W][kRedBlock];
#pragma unroll
    for (int c = 0; c < W; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < W; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < W; ++c) part[blk * S + c] = red[c][0];
    }
}

__device__ inline void vhist_clear(unsigned* h)
{
    for (int b = threadIdx.x; b < 2 * kHistStride; b += blockDim.x) h[b] = 0u;
}

__device__ inline void vhist_flush(const unsigned* h, unsigned long long* g)
{
    for (int b = threadIdx.x; b < 2 * kHistStride; b += blockDim.x)
        if (h[b]) atomicAdd(g + b, (unsigned long long)h[b]);
}

// VDF bin of vverlet.f90:201: INT(mv2/3/dv) + 1 <= Ndist + 1
__device__ inline void vdf_count(unsigned* h, int s, double mv2, double dv)
{
#pragma clang fp contract(off)
    const double t = mv2 / 3.0 / dv;
    if (t >= 0.0 && t < (double)kNbin) atomicAdd(&h[s * kHistStride + (int)t], 1u);
}

// fractional coordinate in [0, 1] -> 32-bit fixed point (modulo 2^32) and the
// FP32 residual of the unwrapped value
__device__ inline void to_fixed(double x, double box, uint32_t* u, float* r)
{
    const double s = rint((x / box + 0.5) * 4294967296.0);
    *u = (uint32_t)(uint64_t)(int64_t)s;
    *r = (float)(x - (s * 0x1p-32 - 0.5) * box);
}

__device__ inline FixPos make_fixpos(double x, double y, double z, const double* box)
{
    FixPos f;
    to_fixed(x, box[0], &f.ux, &f.rx);
    to_fixed(y, box[1], &f.uy, &f.ry);
    to_fixed(z, box[2], &f.uz, &f.rz);
    f.pad0 = 0u;
    f.pad1 = 0.f;
    return f;
}

__device__ inline void vv1_body(const MdArgs& m, int blk, int thermo, double dt, double fric)
{
#pragma clang fp contract(off)
    const int i = blk * kRedBlock + threadIdx.x;
    if (i >= m.n) return;
    const double xm = m.xm[i >= m.ne];
    double* xs[3] = {m.x, m.y, m.z};
    double* vs[3] = {m.vx, m.vy, m.vz};
    const double* fs[3] = {m.fx, m.fy, m.fz};
    double xn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = vs[k][i];
        if (thermo == TCP_STOCH)
            v = v * (1.0 - 0.5 * fric * dt / xm) + 0.5 * dt * fs[k][i] / xm;   // vverlet.f90:37-38
        else
            v = v + 0.5 * dt * fs[k][i] / xm;                                   // vverlet.f90:151
        const double xx = xs[k][i] + dt * v;                                    // vverlet.f90:152-153
        xn[k] = xx - m.box[k] * round(xx / m.box[k]);
        vs[k][i] = v;
        xs[k][i] = xn[k];
    }
    if (m.upos) m.upos[i] = make_fixpos(xn[0], xn[1], xn[2], m.box);
}

__global__ __launch_bounds__(kRedBlock) void vv1_kernel(MdArgs m, int thermo, double dt, double fric)
{
    vv1_body(m, blockIdx.x, thermo, dt, fric);
}

__global__ __launch_bounds__(kRedBlock) void upos_kernel(MdArgs m)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    m.upos[i] = make_fixpos(m.x[i], m.y[i], m.z[i], m.box);
}

// f = sum over chunks of the pair partials (+ reciprocal forces), block
// partials of sum f (3) and of sum u (1).  pu_row > 0 (the Newton kernels'
// block-row width): the energy slots of particle i below its own block row
// i / pu_row hold only zeros and are not written, so they are skipped (the
// sum is the same, bit for bit).
__device__ inline void combine_body(const MdArgs& m, int blk, const double* pf, const double* pu,
                                    int nchunk, int pu_row, const double* frx, const double* fry,
                                    const double* frz, double* part, int accumulate)
{
    const int i = blk * kRedBlock + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < m.n) {
        const size_t n = m.n;
        double f3[3] = {0.0, 0.0, 0.0}, u = 0.0;
        const int c0 = pu_row > 0 ? i / pu_row : 0;  // uniform across a wavefront
        for (int c = 0; c < nchunk; ++c) {
            const double* p = pf + (size_t)c * 3 * n;
            f3[0] += p[i];
            f3[1] += p[n + i];
            f3[2] += p[2 * n + i];
            if (c >= c0) u += pu[(size_t)c * n + i];
        }
        if (accumulate) {
            f3[0] = m.fx[i] + f3[0];
            f3[1] = m.fy[i] + f3[1];
            f3[2] = m.fz[i] + f3[2];
        }
        if (frx) {  // SPME_SUM: ff = ff + grec (force.f90:955)
            f3[0] = f3[0] + frx[i];
            f3[1] = f3[1] + fry[i];
            f3[2] = f3[2] + frz[i];
        }
        m.fx[i] = f3[0];
        m.fy[i] = f3[1];
        m.fz[i] = f3[2];
        v[0] = f3[0]; v[1] = f3[1]; v[2] = f3[2]; v[3] = u;
    }
    block_sum_store<4>(v, part, blk);
}

__global__ __launch_bounds__(kRedBlock) void combine_kernel(MdArgs m, const double* pf,
                                                            const double* pu, int nchunk,
                                                            int pu_row, const double* frx,
                                                            const double* fry, const double* frz,
                                                            double* part, int accumulate)
{
    combine_body(m, blockIdx.x, pf, pu, nchunk, pu_row, frx, fry, frz, part, accumulate);
}

// f += fr (SPME_SUM) after a combine_kernel call that left fr out, and the
// force entries of its block partials (part[b * 4 + 0..2]; entry 3, sum u,
// stays): the same values and partials as a combine_kernel call given fr.
__global__ __launch_bounds__(kRedBlock) void add_recip_kernel(MdArgs m, const double* frx,
                                                              const double* fry,
                                                              const double* frz, double* part)
{
    const int i = blockIdx.x * kRedBlock + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < m.n) {
        v[0] = m.fx[i] + frx[i];
        v[1] = m.fy[i] + fry[i];
        v[2] = m.fz[i] + frz[i];
        m.fx[i] = v[0];
        m.fy[i] = v[1];
        m.fz[i] = v[2];
    }
    block_sum_store<3, 4>(v, part, blockIdx.x);
}

// sums[c] = sum over blocks of part[b * width + c], in block order (one block).
__device__ inline void reduce_parts_body(const double* part, int nblk, int width, double* sums)
{
    __shared__ double red[kRedBlock];
    for (int c = 0; c < width; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += kRedBlock) s += part[(size_t)b * width + c];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = kRedBlock / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) sums[c] = red[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRedBlock) void reduce_parts_kernel(const double* part, int nblk,
                                                                 int width, double* sums)
{
    reduce_parts_body(part, nblk, width, sums);
}

// f += fr (SPME_SUM alone)
__global__ __launch_bounds__(kRedBlock) void add_forces_kernel(MdArgs m, const double* frx,
                                                               const double* fry, const double* frz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    m.fx[i] = m.fx[i] + frx[i];
    m.fy[i] = m.fy[i] + fry[i];
    m.fz[i] = m.fz[i] + frz[i];
}

// Block partials of sum f over all particles
__global__ __launch_bounds__(kRedBlock) void force_sum_kernel(MdArgs m, double* part)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (i < m.n) { v[0] = m.fx[i]; v[1] = m.fy[i]; v[2] = m.fz[i]; }
    block_sum_store<3>(v, part, blockIdx.x);
}

// REMOVE_RIGIDMTN alone: f -= fsum / N (force.f90:996-1004)
__global__ __launch_bounds__(kRedBlock) void sub_mean_kernel(MdArgs m, const double* fsum)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n) return;
    const double N = (double)m.n;
    m.fx[i] = m.fx[i] - fsum[0] / N;
    m.fy[i] = m.fy[i] - fsum[1] / N;
    m.fz[i] = m.fz[i] - fsum[2] / N;
}

// Counter-based unit Gaussian (SplitMix64 hashing of the key, Box-Muller).
__device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ inline double gauss(uint64_t seed, uint64_t counter, uint64_t idx)
{
    const uint64_t h1 = mix64(seed ^ mix64(counter ^ mix64(idx)));
    const uint64_t h2 = mix64(h1 + 0x632be59bd9b4e019ull);
    const double u1 = ((h1 >> 11) + 1) * 0x1p-53;  // (0, 1]
    const double u2 = (h2 >> 11) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Second half kick (vverlet.f90:190-219, 245-253, 319-329, 85-98), with the
// mean-force removal fused in front when fsum != null.  part[blk*2 + s]:
// BEREND/ISOKIN: sum m v_k^2 before rescale; NOTEMP/STOCH: sum m v^2 (final).
__device__ inline void vv2_kick_body(const MdArgs& m, int blk, int thermo, double dt, double Te,
                                     double Ti, double fric, uint64_t seed, uint64_t counter,
                                     const double* fsum, double* part)
{
    __shared__ unsigned hist[2 * kHistStride];
    const bool final_now = thermo == TCP_NOTEMP || thermo == TCP_STOCH;
    if (final_now) {
        vhist_clear(hist);
        __syncthreads();
    }
    const int i = blk * kRedBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (i < m.n) {
#pragma clang fp contract(off)
        const int s = i >= m.ne;
        const double xm = m.xm[s];
        double f[3] = {m.fx[i], m.fy[i], m.fz[i]};
        if (fsum) {
            const double N = (double)m.n;
#pragma unroll
            for (int k = 0; k < 3; ++k) f[k] = f[k] - fsum[k] / N;
        }
        double v[3] = {m.vx[i], m.vy[i], m.vz[i]};
        if (thermo == TCP_STOCH) {
            const double T = s ? Ti : Te;
            const double beta = sqrt(2.0 * fric * T / dt);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double eta = gauss(seed, counter, 3ull * (uint64_t)i + k);
                f[k] = f[k] + eta * beta;
                v[k] = (v[k] + 0.5 * dt * f[k] / xm) / (1.0 + 0.5 * fric * dt / xm);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = v[k] + 0.5 * dt * f[k] / xm;
        }
        m.fx[i] = f[0]; m.fy[i] = f[1]; m.fz[i] = f[2];
        m.vx[i] = v[0]; m.vy[i] = v[1]; m.vz[i] = v[2];
        if (final_now) {
            const double mv2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * xm;
            acc[s] = mv2;
            vdf_count(hist, s, mv2, m.dv[s]);
        } else {
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) a = a + xm * (v[k] * v[k]);
            acc[s] = a;
        }
    }
    block_sum_store<2>(acc, part, blk);
    if (final_now) {
        __syncthreads();
        vhist_flush(hist, m.vdf);
    }
}

__global__ __launch_bounds__(kRedBlock) void vv2_kick_kernel(MdArgs m, int thermo, double dt,
                                                             double Te, double Ti, double fric,
                                                             uint64_t seed, uint64_t counter,
                                                             const double* fsum, double* part)
{
    vv2_kick_body(m, blockIdx.x, thermo, dt, Te, Ti, fric, seed, counter, fsum, part);
}

// lambda per species from the pre-rescale kinetic sums (vverlet.f90:256-258, 330-332).
__device__ inline void thermo_lambda_body(const double* mv2, int thermo, int ne, int ni, double Te,
                                          double Ti, double dt, double tau, double* lam)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    for (int s = 0; s < 2; ++s) {
        const double T = s ? Ti : Te;
        const int ns = s ? ni : ne;
        const double T0 = mv2[s] / (double)(3 * ns);
        lam[s] = thermo == TCP_BEREND ? sqrt(1.0 + dt * (T / T0 - 1.0) / tau) : sqrt(T / T0);
    }
}

__global__ void thermo_lambda_kernel(const double* mv2, int thermo, int ne, int ni, double Te,
                                     double Ti, double dt, double tau, double* lam)
{
    thermo_lambda_body(mv2, thermo, ne, ni, Te, Ti, dt, tau, lam);
}

__device__ inline void vv2_scale_body(const MdArgs& m, int blk, const double* lam, double* part)
{
    __shared__ unsigned hist[2 * kHistStride];
    vhist_clear(hist);
    __syncthreads();
    const int i = blk * kRedBlock + threadIdx.x;
    double acc[2] = {0.0, 0.0};
    if (i < m.n) {
#pragma clang fp contract(off)
        const int s = i >= m.ne;
        const double xm = m.xm[s], l = lam[s];
        const double v0 = m.vx[i] * l, v1 = m.vy[i] * l, v2 = m.vz[i] * l;
        m.vx[i] = v0; m.vy[i] = v1; m.vz[i] = v2;
        const double mv2 = (v0 * v0 + v1 * v1 + v2 * v2) * xm;
        acc[s] = mv2;
        vdf_count(hist, s, mv2, m.dv[s]);
    }
    block_sum_store<2>(acc, part, blk);
    __syncthreads();
    vhist_flush(hist, m.vdf);
}

__global__ __launch_bounds__(kRedBlock) void vv2_scale_kernel(MdArgs m, const double* lam,
                                                              double* part)
{
    vv2_scale_body(m, blockIdx.x, lam, part);
}

// One row of the per-step trace: Ur, Uo, Um, Epot = (Ur + Uo) eps, mv2el, mv2ion.
__device__ inline void trace_body(double* row, const double* usum, double ur_scale, double Uo,
                                  const double* um, const double* mv2, double eps)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    const double Ur = usum[0] * ur_scale;
    row[0] = Ur;
    row[1] = Uo;
    row[2] = um ? um[0] : 0.0;
    row[3] = (Ur + Uo) * eps;
    row[4] = mv2[0];
    row[5] = mv2[1];
}

__global__ void trace_kernel(double* row, const double* usum, double ur_scale, double Uo,
                             const double* um, const double* mv2, double eps)
{
    trace_body(row, usum, ur_scale, Uo, um, mv2, eps);
}

// Um = eps/(2 pi V) * sum of the energy partials, one block, fixed order.
__device__ inline void um_body(const double* epart, int n, double scale, double* um)
{
    __shared__ double red[kRedBlock];
    double s = 0.0;
    for (int b = threadIdx.x; b < n; b += kRedBlock) s += epart[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kRedBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) um[0] = red[0] * scale;
}

__global__ __launch_bounds__(kRedBlock) void um_kernel(const double* epart, int n, double scale,
                                                       double* um)
{
    um_body(epart, n, scale, um);
}

// ------------------------------------------------------------- batches ---
//
// Many independent replicas per launch: blockIdx.y selects the replica, whose
// arguments live in a device table, and blockIdx.x is the block index the
// single-replica kernel would have had (blocks past the replica's size exit
// at once).  The bodies are the ones above, so every replica computes exactly
// what a launch for it alone computes.

__device__ inline bool batch_block_live(const BatchRep& r)
{
    return (int)blockIdx.x < md_blocks_dev(r.m.n);
}

__global__ __launch_bounds__(kRedBlock) void vv1_batch_kernel(const BatchRep* reps, int thermo,
                                                              double dt, double fric)
{
    const BatchRep& r = reps[blockIdx.y];
    if (!batch_block_live(r)) return;
    vv1_body(r.m, blockIdx.x, thermo, dt, fric);
}

__global__ __launch_bounds__(kRedBlock) void combine_batch_kernel(const BatchRep* reps)
{
    const BatchRep& r = reps[blockIdx.y];
    if (!batch_block_live(r)) return;
    const bool rec = r.recip != TCP_RECIP_NONE;
    combine_body(r.m, blockIdx.x, r.pa.pf, r.pa.pu, r.slots, r.pu_row, rec ? r.ra.frx : nullptr,
                 rec ? r.ra.fry : nullptr, rec ? r.ra.frz : nullptr, r.part, 0);
}

__global__ __launch_bounds__(kRedBlock) void reduce_parts_batch_kernel(const BatchRep* reps,
                                                                       int width, int slot)
{
    const BatchRep& r = reps[blockIdx.x];
    reduce_parts_body(r.part, md_blocks_dev(r.m.n), width, r.sc + slot);
}

__global__ __launch_bounds__(kRedBlock) void vv2_kick_batch_kernel(const BatchRep* reps, int thermo,
                                                                   double dt, double fric,
                                                                   uint64_t seed,
                                                                   const uint64_t* counter)
{
    const BatchRep& r = reps[blockIdx.y];
    if (!batch_block_live(r)) return;
    vv2_kick_body(r.m, blockIdx.x, thermo, dt, r.Te, r.Ti, fric, seed, counter[0],
                  r.sc + SC_FSUM, r.part);
}

__global__ void thermo_lambda_batch_kernel(const BatchRep* reps, int thermo, double dt, double tau)
{
    const BatchRep& r = reps[blockIdx.x];
    thermo_lambda_body(r.sc + SC_MV2, thermo, r.m.ne, r.m.n - r.m.ne, r.Te, r.Ti, dt, tau,
                       r.sc + SC_LAM);
}

__global__ __launch_bounds__(kRedBlock) void vv2_scale_batch_kernel(const BatchRep* reps)
{
    const BatchRep& r = reps[blockIdx.y];
    if (!batch_block_live(r)) return;
    vv2_scale_body(r.m, blockIdx.x, r.sc + SC_LAM, r.part);
}

__global__ void trace_batch_kernel(const BatchRep* reps, int nrep, double* trace,
                                   const uint64_t* counter)
{
    const BatchRep& r = reps[blockIdx.x];
    double* row = trace + 6 * ((size_t)counter[1] * nrep + blockIdx.x);
    trace_body(row, r.sc + SC_USUM, r.ur_scale, r.Uo,
               r.recip != TCP_RECIP_NONE ? r.sc + SC_UM : nullptr, r.sc + SC_MV2, r.eps);
}

__global__ __launch_bounds__(kRedBlock) void um_batch_kernel(const BatchRep* reps)
{
    const BatchRep& r = reps[blockIdx.x];
    if (r.recip == TCP_RECIP_NONE) return;
    // the k-sum leaves one energy partial, SPME one per z-pass block
    um_body(r.ra.epart, r.recip == TCP_RECIP_SPME ? r.ra.nepart : 1, r.um_scale, r.sc + SC_UM);
}

// counter[0]: step counter of the Gaussian noise, counter[1]: trace row
__global__ void step_advance_kernel(uint64_t* counter)
{
    if (threadIdx.x == 0) {
        counter[0] += 1;
        counter[1] += 1;
    }
}

// Host-state transposes: one thread per particle, the AoS triple read or
// written as three consecutive doubles (the block's lanes cover a contiguous
// 24 B x 256 span).
__global__ __launch_bounds__(256) void aos_to_soa_kernel(const double* __restrict__ aos,
                                                         double* __restrict__ x,
                                                         double* __restrict__ y,
                                                         double* __restrict__ z, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    x[i] = aos[3 * (size_t)i];
    y[i] = aos[3 * (size_t)i + 1];
    z[i] = aos[3 * (size_t)i + 2];
}

__global__ __launch_bounds__(256) void soa_to_aos_kernel(const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ z,
                                                         double* __restrict__ aos, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    aos[3 * (size_t)i] = x[i];
    aos[3 * (size_t)i + 1] = y[i];
    aos[3 * (size_t)i + 2] = z[i];
}

}  // namespace

int md_blocks(int n) { return (n + kRedBlock - 1) / kRedBlock; }

void launch_vv1(const MdArgs& m, int thermo, double dt, double fric, hipStream_t s)
{
    hipLaunchKernelGGL(vv1_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, thermo, dt, fric);
}

void launch_aos_to_soa(const double* aos, double* x, double* y, double* z, int n, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(aos_to_soa_kernel, dim3((n + 255) / 256), dim3(256), 0, s, aos, x, y, z, n);
}

void launch_soa_to_aos(const double* x, const double* y, const double* z, double* aos, int n,
                       hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(soa_to_aos_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, y, z, aos, n);
}

void launch_make_upos(const MdArgs& m, hipStream_t s)
{
    hipLaunchKernelGGL(upos_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m);
}

void launch_combine(const MdArgs& m, const double* pf, const double* pu, int nchunk, int pu_row,
                    const double* frx, const double* fry, const double* frz, double* part,
                    bool accumulate, hipStream_t s)
{
    hipLaunchKernelGGL(combine_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, pf, pu,
                       nchunk, pu_row, frx, fry, frz, part, (int)accumulate);
}

void launch_add_recip(const MdArgs& m, const double* frx, const double* fry, const double* frz,
                      double* part, hipStream_t s)
{
    hipLaunchKernelGGL(add_recip_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, frx, fry,
                       frz, part);
}

void launch_reduce_parts(const double* part, int nblk, int width, double* sums, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(kRedBlock), 0, s, part, nblk, width, sums);
}

void launch_add_forces(const MdArgs& m, const double* frx, const double* fry, const double* frz,
                       hipStream_t s)
{
    hipLaunchKernelGGL(add_forces_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, frx, fry,
                       frz);
}

void launch_force_sum(const MdArgs& m, double* part, hipStream_t s)
{
    hipLaunchKernelGGL(force_sum_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, part);
}

void launch_sub_mean(const MdArgs& m, const double* fsum, hipStream_t s)
{
    hipLaunchKernelGGL(sub_mean_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, fsum);
}

void launch_vv2_kick(const MdArgs& m, int thermo, double dt, double Te, double Ti, double fric,
                     uint64_t seed, uint64_t counter, const double* fsum, double* part,
                     hipStream_t s)
{
    hipLaunchKernelGGL(vv2_kick_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, thermo, dt,
                       Te, Ti, fric, seed, counter, fsum, part);
}

void launch_thermo_lambda(const double* mv2, int thermo, int ne, int ni, double Te, double Ti,
                          double dt, double tau, double* lam, hipStream_t s)
{
    hipLaunchKernelGGL(thermo_lambda_kernel, dim3(1), dim3(64), 0, s, mv2, thermo, ne, ni, Te, Ti,
                       dt, tau, lam);
}

void launch_vv2_scale(const MdArgs& m, const double* lam, double* part, hipStream_t s)
{
    hipLaunchKernelGGL(vv2_scale_kernel, dim3(md_blocks(m.n)), dim3(kRedBlock), 0, s, m, lam, part);
}

void launch_trace(double* row, const double* usum, double ur_scale, double Uo, const double* um,
                  const double* mv2, double eps, hipStream_t s)
{
    hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(64), 0, s, row, usum, ur_scale, Uo, um, mv2, eps);
}

void launch_um(const double* epart, int n, double scale, double* um, hipStream_t s)
{
    hipLaunchKernelGGL(um_kernel, dim3(1), dim3(kRedBlock), 0, s, epart, n, scale, um);
}

void launch_vv1_batch(const BatchRep* reps, int nrep, int maxblk, int thermo, double dt,
                      double fric, hipStream_t s)
{
    hipLaunchKernelGGL(vv1_batch_kernel, dim3(maxblk, nrep), dim3(kRedBlock), 0, s, reps, thermo,
                       dt, fric);
}

void launch_combine_batch(const BatchRep* reps, int nrep, int maxblk, hipStream_t s)
{
    hipLaunchKernelGGL(combine_batch_kernel, dim3(maxblk, nrep), dim3(kRedBlock), 0, s, reps);
}

void launch_reduce_parts_batch(const BatchRep* reps, int nrep, int width, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(reduce_parts_batch_kernel, dim3(nrep), dim3(kRedBlock), 0, s, reps, width,
                       slot);
}

void launch_vv2_batch(const BatchRep* reps, int nrep, int maxblk, int thermo, double dt, double tau,
                      double fric, uint64_t seed, const uint64_t* counter, hipStream_t s)
{
    hipLaunchKernelGGL(vv2_kick_batch_kernel, dim3(maxblk, nrep), dim3(kRedBlock), 0, s, reps,
                       thermo, dt, fric, seed, counter);
    if (thermo == TCP_BEREND || thermo == TCP_ISOKIN) {
        hipLaunchKernelGGL(reduce_parts_batch_kernel, dim3(nrep), dim3(kRedBlock), 0, s, reps, 2,
                           (int)SC_MV2);
        hipLaunchKernelGGL(thermo_lambda_batch_kernel, dim3(nrep), dim3(64), 0, s, reps, thermo, dt,
                           tau);
        hipLaunchKernelGGL(vv2_scale_batch_kernel, dim3(maxblk, nrep), dim3(kRedBlock), 0, s, reps);
    }
    hipLaunchKernelGGL(reduce_parts_batch_kernel, dim3(nrep), dim3(kRedBlock), 0, s, reps, 2,
                       (int)SC_MV2);
}

void launch_um_batch(const BatchRep* reps, int nrep, hipStream_t s)
{
    hipLaunchKernelGGL(um_batch_kernel, dim3(nrep), dim3(kRedBlock), 0, s, reps);
}

void launch_trace_batch(const BatchRep* reps, int nrep, double* trace, const uint64_t* counter,
                        hipStream_t s)
{
    hipLaunchKernelGGL(trace_batch_kernel, dim3(nrep), dim3(64), 0, s, reps, nrep, trace, counter);
}

void launch_step_advance(uint64_t* counter, hipStream_t s)
{
    hipLaunchKernelGGL(step_advance_kernel, dim3(1), dim3(64), 0, s, counter);
}

}  // namespace tcp

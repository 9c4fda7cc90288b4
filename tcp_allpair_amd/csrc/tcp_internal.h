// Internal declarations shared by the HIP translation units of libtcp_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tcp_gpu.h"

namespace tcp {

constexpr int kNbin = TCP_NBIN;        // Ndist+1
constexpr int kHistStride = 512;       // padded bins per histogram in device memory
constexpr int kPairBlock = 256;        // threads (i particles) per pair-kernel block
constexpr int kRedBlock = 256;         // threads per block of the O(N) kernels
constexpr int64_t kMaxParticles = int64_t(1) << 22;  // per context (block-row work items stay in int)
constexpr double kMixedErfcMax = 4.5;  // the mixed kernels' erfcx table covers x = alpha r up to it

// Fixed-point position for the mixed pair kernel: u = round((x/L + 1/2) 2^32)
// modulo 2^32 per axis and the FP32 residual r = x - (u 2^-32 - 1/2) L of the
// unwrapped u.  The wrapped 32-bit difference of two u is the minimum-image
// separation in units of L 2^-32 exactly, and adding the residual difference
// restores the FP64 separation to FP32 rounding: dx = float(int(u_i - u_j)) s + (r_i - r_j).
struct __align__(32) FixPos {
    uint32_t ux, uy, uz, pad0;
    float rx, ry, rz, pad1;
};

// Particle positions as structure of arrays, electrons first.
struct PosSoA {
    const double* x;
    const double* y;
    const double* z;
};

// Arguments of the all-pair real-space kernels.  The i-set and the j-set are
// both "electrons first"; i == j pairs are skipped through r2 > 0.
struct PairArgs {
    PosSoA pi;                 // i-set (FP64 path)
    PosSoA pj;                 // j-set (FP64 path)
    const FixPos* ui;          // i-set fixed-point coordinates (mixed path)
    const FixPos* uj;          // j-set fixed-point coordinates (mixed path)
    int n_i, ne_i;             // i-set size and electron count
    int n_j, ne_j;             // j-set size and electron count
    int jchunk;                // j indices per blockIdx.y
    double box[3], inv_box[3];
    double rc2, alpha, rs, eps;
    double q[2];
    double inv_dx[3];          // RDF bin widths ee, ei, ii (reciprocal, mixed path)
    double dx[3];              // RDF bin widths (FP64 path divides as the reference)
    unsigned rdf_weight;       // 0: no RDF, 1: ordered pairs, 2: unordered pairs
    int flags;                 // TCP_KELBG | TCP_REG_EE
    unsigned long long* rdf;   // [3][kHistStride] device counts
    double* pf;                // [nchunk][3][n_i] partial forces (eps included)
    double* pu;                // [nchunk][n_i] partial energies sum_j q q (erfc - y)/r
};

void launch_pair_f64(const PairArgs& a, int nchunk, hipStream_t s);
void launch_pair_mixed(const PairArgs& a, int nchunk, hipStream_t s);
// Self interaction with Newton's third law (both precisions): block-row slots
// of partials, pf[rows][3][n] and pu[rows][n] with rows =
// n3_block_rows_shape(n, shape); blkmap lists the (row, column) block pairs
// with row <= column.
// persistent launch of grid resident blocks over the nblk block pairs of blkmap;
// queue[2] is zero before the first launch and left zero by every launch
// shape: 0 (8 x 8 tile blocks) or kFineN3Shape (4 x 4 tile blocks, small systems)
int pair_n3_occupancy(bool rdf, bool kelbg, int shape);
void launch_pair_mixed_n3(const PairArgs& a, const int2* blkmap, int nblk, int* queue, int grid,
                          int shape, hipStream_t s);
int pair_f64_n3_occupancy(bool rdf, bool kelbg, int shape);
void launch_pair_f64_n3(const PairArgs& a, const int2* blkmap, int nblk, int* queue, int grid,
                        int shape, hipStream_t s);

// SPME / Ewald reciprocal part
struct RecipArgs {
    int n, ne;
    const double* x;
    const double* y;
    const double* z;
    int K[3];
    int order;
    double box[3];
    double q[2];
    const double* qp;          // per-particle charges (null: the species charges q)
    double* pot;               // [n] dUm/dq_m when not null (gather and k-sum kernels)
    double eps;
    double* Q;                 // real charge grid K1*K2*K3 (x fastest)
    double2* A;                // complex work grid
    double* C;                 // real convolved grid
    const double* gb;          // PSI * B modulus, K1*K2*K3
    const double2* tw[3];      // twiddles exp(+2 pi i m/K_d), m < K_d/2, per axis
    double* epart;             // energy partials (one per block of the convolution)
    int nepart;
    double* frx;               // reciprocal forces (SoA)
    double* fry;
    double* frz;
    double2* skpart;           // direct k-sum: per-chunk structure-factor partials
    // brick-binned spreading (used when every K_d is a multiple of kBrick)
    int nbrick[3];             // bricks per axis (0: unbinned atomic spreading)
    int* brick_of;             // [n] brick of each particle's first grid point
    int* brick_count;          // [nbrick]
    int* brick_start;          // [nbrick + 1]
    int* brick_cursor;         // [nbrick]
    int* sorted;               // [n] particle indices grouped by brick
    double* tiles;             // [nbrick][(kBrick + P - 1)^3] per-brick grid tiles
};

constexpr int kBrick = 8;      // grid points per brick edge

void launch_spme_spread(const RecipArgs& a, hipStream_t s);
// Q -> C = Q convolved with IDFT(gb), energy partials gb |S|^2 per block into epart
void launch_spme_convolve(const RecipArgs& a, hipStream_t s);
void launch_spme_gather(const RecipArgs& a, hipStream_t s);
void launch_fill(double* p, size_t n, double v, hipStream_t s);
// AoS (x0 y0 z0 x1 ...) <-> SoA transposes of n particles
void launch_aos_to_soa(const double* aos, double* x, double* y, double* z, int n, hipStream_t s);
void launch_soa_to_aos(const double* x, const double* y, const double* z, double* aos, int n,
                       hipStream_t s);
int fft_epart_blocks(const int* K);

// Direct Ewald k-space sum (small systems, reference-independent extension):
// sk_work holds 3 (2 kmax + 1)^3 doubles, a.skpart ewald_half(kmax) x
// ceil(n / ewald_chunk(kmax)) structure-factor partials; energy sum into epart[0].
int ewald_chunk(int kmax);
int ewald_half(int kmax);
void launch_ewald_kspace(const RecipArgs& a, int kmax, double alpha, double* sk_work,
                         hipStream_t s);

// Energy derivatives (FP64, ordered pairs, per-particle charges): per chunk of j
// and particle i, part[(chunk * kAdjQ + k) * n + i] holds k = 0..2 the force
// (eps included), 3 sum_j q_i q_j (erfc + R)/r, 4 phi_i = sum_j q_j (erfc + R)/r,
// 5 sum_j q_i q_j d(R/r)/dlambda, 6..8 dF_i/dlambda (eps included).
constexpr int kAdjQ = 9;
struct AdjArgs {
    PosSoA p;
    const double* qp;
    int n, ne, jchunk, flags;
    double box[3], inv_box[3];
    double rc2, alpha, rs, eps;
    double* part;              // [nchunk][kAdjQ][n]
};
void launch_pair_adjoint(const AdjArgs& a, int nchunk, hipStream_t s);
// out[k * n + i] = sum over chunks of part, chunk order
void launch_adjoint_reduce(const double* part, int nchunk, int n, double* out, hipStream_t s);

// Integrator / reductions
struct MdArgs {
    int n, ne;
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    double* fx; double* fy; double* fz;
    FixPos* upos;              // fixed-point mirror of x, y, z
    double box[3];
    double xm[2];
    double dv[2];              // VDF bin widths
    unsigned long long* vdf;   // [2][kHistStride]
};

int md_blocks(int n);
__host__ __device__ inline int md_blocks_dev(int n) { return (n + kRedBlock - 1) / kRedBlock; }
void launch_vv1(const MdArgs& m, int thermo, double dt, double fric, hipStream_t s);
void launch_make_upos(const MdArgs& m, hipStream_t s);
// f (=|+=) sum over chunks of the pair partials (+ fr when frx != null);
// block partials (sum f, sum u) into part[4 * md_blocks(n)].  pu_row: block-row
// width of the Newton kernels, whose energy slots below a particle's own row
// are not written (0: every slot is read).
void launch_combine(const MdArgs& m, const double* pf, const double* pu, int nchunk, int pu_row,
                    const double* frx, const double* fry, const double* frz, double* part,
                    bool accumulate, hipStream_t s);
// f += fr and the force entries of combine_kernel's block partials, for a
// combine_kernel call that left fr out (fr computed beside it)
void launch_add_recip(const MdArgs& m, const double* frx, const double* fry, const double* frz,
                      double* part, hipStream_t s);
// sums[c] = sum_b part[b * width + c] in block order (single block)
void launch_reduce_parts(const double* part, int nblk, int width, double* sums, hipStream_t s);
void launch_add_forces(const MdArgs& m, const double* frx, const double* fry, const double* frz,
                       hipStream_t s);
void launch_force_sum(const MdArgs& m, double* part, hipStream_t s);
void launch_sub_mean(const MdArgs& m, const double* fsum, hipStream_t s);
void launch_vv2_kick(const MdArgs& m, int thermo, double dt, double Te, double Ti, double fric,
                     uint64_t seed, uint64_t counter, const double* fsum, double* part,
                     hipStream_t s);
void launch_thermo_lambda(const double* mv2, int thermo, int ne, int ni, double Te, double Ti,
                          double dt, double tau, double* lam, hipStream_t s);
void launch_vv2_scale(const MdArgs& m, const double* lam, double* part, hipStream_t s);
void launch_trace(double* row, const double* usum, double ur_scale, double Uo, const double* um,
                  const double* mv2, double eps, hipStream_t s);
void launch_um(const double* epart, int n, double scale, double* um, hipStream_t s);

// Device scalars of a context (double slots)
enum {
    SC_FSUM = 0,   // 3: sum of forces
    SC_USUM = 3,   // sum_i u_i
    SC_UM = 4,     // Um
    SC_MV2 = 5,    // 2: mv2el, mv2ion
    SC_LAM = 7,    // 2: thermostat lambdas
    SC_N = 16
};

// ------------------------------------------------------------- batches ---
// One replica of a batch of independent plasmas: the arguments every batched
// kernel needs, kept in a device table indexed by replica.
struct BatchRep {
    MdArgs m;
    PairArgs pa;               // pair kernel arguments; pf/pu hold `slots` partial slots
    RecipArgs ra;              // reciprocal-part arguments (k-sum or SPME)
    double* sk;                // k-sum work, 3 (2 kmax + 1)^3
    double* part;              // block partials
    double* sc;                // device scalars SC_*
    int slots;                 // pair partial slots summed by combine
    int pu_row;                // particles per block row of the pair kernel's shape
    int recip;                 // TCP_RECIP_NONE, _EWALD or _SPME (ra holds its buffers)
    int kmax;
    double alpha;
    double Te, Ti;             // thermostat targets
    double ur_scale, Uo, eps, um_scale;
};

// Work item of a persistent pair launch: replica, then (row, column) block pair
// (Newton's-third-law kernel) or (i block, j chunk) (FP64 kernel).
struct PairItem {
    int rep, a, b, cost;
};

void launch_vv1_batch(const BatchRep* reps, int nrep, int maxblk, int thermo, double dt,
                      double fric, hipStream_t s);
void launch_combine_batch(const BatchRep* reps, int nrep, int maxblk, hipStream_t s);
void launch_reduce_parts_batch(const BatchRep* reps, int nrep, int width, int slot, hipStream_t s);
// second half kick (+ thermostat) and the kinetic sums into sc[SC_MV2]
void launch_vv2_batch(const BatchRep* reps, int nrep, int maxblk, int thermo, double dt, double tau,
                      double fric, uint64_t seed, const uint64_t* counter, hipStream_t s);
void launch_um_batch(const BatchRep* reps, int nrep, hipStream_t s);
void launch_trace_batch(const BatchRep* reps, int nrep, double* trace, const uint64_t* counter,
                        hipStream_t s);
void launch_step_advance(uint64_t* counter, hipStream_t s);
// direct k-space sums of every replica (same kmax): S(m), forces, energy partial;
// maxn: particles of the largest replica
void launch_ewald_batch(const BatchRep* reps, int nrep, int kmax, int maxn, hipStream_t s);
// SPME of every replica (brick-binned grids, one spline order): the launch
// sizes and a device table of each replica's five FFT passes
constexpr int kSpmeBatchPasses = 5;
struct SpmeBatchPlan {
    void* passes = nullptr;  // device FftPass[kSpmeBatchPasses][nrep]
    int nrep = 0, order = 0, maxn = 0, maxnb = 0, maxgblk = 0;
    int pass_blocks[kSpmeBatchPasses] = {};
};
// ra: host copies of the replicas' RecipArgs; 0 or TCP_ERR_HIP
int spme_batch_plan(const RecipArgs* ra, int nrep, SpmeBatchPlan* plan);
void spme_batch_release(SpmeBatchPlan* plan);
// spreading, convolution and gather of every replica, then Um (sc[SC_UM])
void launch_spme_batch(const BatchRep* reps, const SpmeBatchPlan& plan, hipStream_t s);

// Pair-kernel shapes of the persistent Newton's-third-law kernel: tiles per
// block side and wavefronts per block.  Shape 0 is the single-context kernel's.
constexpr int kNumN3Shapes = 6;
struct N3Shape {
    int tiles, waves;
};
N3Shape n3_shape(int shape);
int n3_block_rows_shape(int n, int shape);
// the 4 x 4 tile, four-wave shape that a single context takes when the 8 x 8
// tile blocks are fewer than four times its resident blocks: their few rounds
// of long items would leave the chip idle at the end
constexpr int kFineN3Shape = 3;
// that rule, from the particle count and the resident blocks of shape 0
inline int n3_self_shape(int n, int grid0)
{
    const int64_t rows0 = n3_block_rows_shape(n, 0);
    return rows0 * (rows0 + 1) / 2 < 4 * (int64_t)grid0 ? kFineN3Shape : 0;
}
// Persistent launches over a work list; queue[2] must be zero before the first
// launch and is left zero by every launch.  grid: resident blocks to start.
void launch_pair_mixed_n3_batch(const BatchRep* reps, const PairItem* items, int nitems,
                                int* queue, int shape, bool rdf, bool kelbg, int grid,
                                hipStream_t s);
// FP64: shape 0 or kFineN3Shape, the block shape of the items' replicas
void launch_pair_f64_batch(const BatchRep* reps, const PairItem* items, int nitems, int* queue,
                           bool rdf, bool kelbg, int grid, int shape, hipStream_t s);
// resident blocks per CU of the persistent kernels
int pair_batch_occupancy(bool mixed, int shape, bool rdf, bool kelbg);

}  // namespace tcp

// tcp_force_direct_noalloc: the real-space all-pair forces of tcp_force_direct
// (Force_direct_self, force.f90:52-180) for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).  The
// caller passes device inputs and outputs; every other buffer the kernels use
// (fixed-point positions, block-row partial slots, block partials, the work
// queue and block map of the persistent pair launch, device scalars) lives in
// one file-scope device array, `scratch`, sized for TCP_NOALLOC_MAX_N
// particles.  Each GPU holds its own copy of the array from the module's load
// to the end of the process; its device address is looked up once per device
// with hipGetSymbolAddress and only that address is handed to kernels.
//
// All calls on one GPU share its array: calls that overlap on one GPU (two
// streams, or two host threads) race.  Calls on one stream are ordered.
#include <mutex>
#include <string>

#include "tcp_ctx.h"

using namespace tcp;

namespace {

constexpr int kNaMaxN = TCP_NOALLOC_MAX_N;
// block rows of the smallest block shape (4 x 4 tiles, 256 particles per row)
constexpr int kNaMaxRows = (kNaMaxN + 255) / 256;
constexpr int kNaMaxBlocks = kNaMaxRows * (kNaMaxRows + 1) / 2;
constexpr int kNaMaxRed = (kNaMaxN + kRedBlock - 1) / kRedBlock;

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// byte offsets of the regions of scratch
constexpr size_t kOffQueue = 0;                                                   // int[2]
constexpr size_t kOffSc = kOffQueue + 256;                                        // double[SC_N]
constexpr size_t kOffPart = kOffSc + align256(SC_N * sizeof(double));             // double[4][red]
constexpr size_t kOffMap = kOffPart + align256(4 * kNaMaxRed * sizeof(double));   // int2[blocks]
constexpr size_t kOffUpos = kOffMap + align256(kNaMaxBlocks * sizeof(int2));      // FixPos[N]
constexpr size_t kOffPu = kOffUpos + align256(kNaMaxN * sizeof(FixPos));          // double[rows][N]
constexpr size_t kOffPf = kOffPu + align256((size_t)kNaMaxRows * kNaMaxN * sizeof(double));
constexpr size_t kScratchBytes = kOffPf + align256((size_t)kNaMaxRows * 3 * kNaMaxN * sizeof(double));
static_assert(kScratchBytes < (size_t)16 << 20, "scratch stays below 16 MiB");

}  // namespace

// the array itself; host code never passes this name to a kernel, only the
// device address that hipGetSymbolAddress returns for it
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kMaxDevices];  // device address of scratch per device (null: not looked up)
int g_na_cus[kMaxDevices];

// Looks up the current device's copy of scratch once (with that device current).
int na_setup(int* dev, unsigned char** base, int* cus, std::string* msg)
{
    if (hipGetDevice(dev) != hipSuccess || *dev < 0 || *dev >= kMaxDevices) {
        *msg = "no HIP device";
        return TCP_ERR_HIP;
    }
    std::lock_guard<std::mutex> l(g_na_mu);
    if (!g_na_ptr[*dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return TCP_ERR_HIP;
        }
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, *dev) != hipSuccess || n < 1) {
            *msg = "no multiprocessor count";
            return TCP_ERR_HIP;
        }
        g_na_cus[*dev] = n;
        g_na_ptr[*dev] = p;
    }
    *base = (unsigned char*)g_na_ptr[*dev];
    *cus = g_na_cus[*dev];
    return 0;
}

// Zeroes the work queue and writes the block map of the persistent pair
// launch: the off-diagonal block pairs (r, q), r < q, row by row, then the
// diagonal ones (the order tcp_create uses).  One thread per block row.
__global__ void na_prepare_kernel(int* queue, int2* map, int rows)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0) {
        queue[0] = 0;
        queue[1] = 0;
    }
    if (r >= rows) return;
    int k = r * (rows - 1) - r * (r - 1) / 2;  // off-diagonal pairs of rows before r
    for (int q = r + 1; q < rows; ++q) map[k++] = make_int2(r, q);
    const int noff = rows * (rows - 1) / 2;
    map[noff + r] = make_int2(r, r);
}

__global__ void na_finish_kernel(const double* sc, double* Ur)
{
    if (threadIdx.x == 0) Ur[0] = sc[SC_USUM];
}

}  // namespace

extern "C" int tcp_force_direct_noalloc(const tcp_params* p, int precision, const double* x,
                                        const double* y, const double* z, double* fx, double* fy,
                                        double* fz, double* Ur, unsigned long long* rdf, void* stream)
{
    std::string msg;
    if (!p || !x || !y || !z || !fx || !fy || !fz || !Ur)
        return set_create_error(TCP_ERR_ARG, "null argument");
    {
        const int rc = validate_params(p, precision, &msg);
        if (rc) return set_create_error(rc, msg.c_str());
    }
    const int n = p->ne + p->ni;
    if (n > kNaMaxN)
        return set_create_error(TCP_ERR_ARG, "tcp_force_direct_noalloc takes at most "
                                                 "TCP_NOALLOC_MAX_N (8192) particles");
    int dev = 0, cus = 1;
    unsigned char* base = nullptr;
    {
        const int rc = na_setup(&dev, &base, &cus, &msg);
        if (rc) return set_create_error(rc, msg.c_str());
    }
    hipStream_t s = (hipStream_t)stream;
    const bool mixed = precision == TCP_MIXED;
    const bool kel = (p->flags & TCP_KELBG) != 0;
    // the block shape and resident blocks tcp_create picks for this system
    const int shape = n3_self_shape(n, cus * (mixed ? pair_n3_occupancy(true, kel, 0)
                                                    : pair_f64_n3_occupancy(true, kel, 0)));
    const int grid = cus * (mixed ? pair_n3_occupancy(rdf != nullptr, kel, shape)
                                  : pair_f64_n3_occupancy(rdf != nullptr, kel, shape));
    const int rows = n3_block_rows_shape(n, shape);
    const int nblk = rows * (rows + 1) / 2;
    if (rows > kNaMaxRows || nblk > kNaMaxBlocks)
        return set_create_error(TCP_ERR_ARG, "block rows exceed the scratch layout");

    int* queue = (int*)(base + kOffQueue);
    double* sc = (double*)(base + kOffSc);
    double* part = (double*)(base + kOffPart);
    int2* map = (int2*)(base + kOffMap);
    FixPos* upos = (FixPos*)(base + kOffUpos);
    double* pu = (double*)(base + kOffPu);
    double* pf = (double*)(base + kOffPf);

    MdArgs m{};
    m.n = n;
    m.ne = p->ne;
    m.x = const_cast<double*>(x);
    m.y = const_cast<double*>(y);
    m.z = const_cast<double*>(z);
    m.fx = fx;
    m.fy = fy;
    m.fz = fz;
    m.upos = upos;
    for (int d = 0; d < 3; ++d) m.box[d] = p->box[d];
    m.xm[0] = p->xm[0];
    m.xm[1] = p->xm[1];

    PairArgs a{};
    for (int d = 0; d < 3; ++d) {
        a.box[d] = p->box[d];
        a.inv_box[d] = 1.0 / p->box[d];
    }
    a.alpha = p->alpha > 0 ? p->alpha : 3.1 / p->rc;  // force.f90:23
    a.rc2 = p->rc * p->rc;
    a.rs = p->rs;
    a.eps = p->eps;
    a.q[0] = p->q[0];
    a.q[1] = p->q[1];
    const double dx[3] = {p->dx_ee, p->dx_ei, p->dx_ii};
    for (int t = 0; t < 3; ++t) {
        a.dx[t] = dx[t];
        a.inv_dx[t] = 1.0 / dx[t];
    }
    a.flags = p->flags;
    a.rdf = rdf;
    a.rdf_weight = rdf ? 1u : 0u;
    a.pi = PosSoA{x, y, z};
    a.pj = a.pi;
    a.ui = upos;
    a.uj = upos;
    a.n_i = a.n_j = n;
    a.ne_i = a.ne_j = p->ne;
    a.jchunk = n;
    a.pf = pf;
    a.pu = pu;

    hipLaunchKernelGGL(na_prepare_kernel, dim3((rows + 63) / 64), dim3(64), 0, s, queue, map, rows);
    if (mixed) launch_make_upos(m, s);
    if (mixed) launch_pair_mixed_n3(a, map, nblk, queue, grid, shape, s);
    else launch_pair_f64_n3(a, map, nblk, queue, grid, shape, s);
    launch_combine(m, pf, pu, rows, n3_shape(shape).tiles * 64, nullptr, nullptr, nullptr, part,
                   false, s);
    launch_reduce_parts(part, md_blocks(n), 4, sc + SC_FSUM, s);
    hipLaunchKernelGGL(na_finish_kernel, dim3(1), dim3(64), 0, s, sc, Ur);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_create_error(TCP_ERR_HIP, hipGetErrorString(e));
    return 0;
}

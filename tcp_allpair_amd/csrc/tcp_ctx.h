// The context object behind the C ABI (tcp_gpu.hip) and the argument
// builders the batch code (batch.hip) shares with it.
#pragma once

#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tcp_internal.h"

namespace tcp {
constexpr int kEvPool = 64;  // timed pair launches kept per call
constexpr int kStageChunks = 16;  // pipelined pieces of one host transfer

// Host memcpy split over the calling thread and a few resident workers (the
// copies between the caller's arrays and the pinned staging buffer).  One pool
// serves the process, created on first use and never torn down, so no exit
// path waits on its threads; copy() calls from several host threads take
// turns.  The workers only copy memory: no HIP or MPI call runs on them.  In a
// child process made by fork() (which has no copy of the workers) copy() is a
// plain memcpy on the calling thread.
class CopyPool {
public:
    static CopyPool& shared();
    void copy(void* dst, const void* src, size_t bytes);

private:
    explicit CopyPool(int workers);
    void work(int id);
    const int parts_;
    const int pid_;  // process that started the workers
    std::vector<std::thread> th_;
    std::mutex job_;  // one copy() at a time
    std::mutex mu_;
    std::condition_variable go_, done_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
    unsigned gen_ = 0;
    int left_ = 0;
};
}  // namespace tcp

struct tcp_ctx {
    tcp_params p{};
    int precision = TCP_FP64;
    int N = 0, Ne = 0;
    int device = 0;
    double alpha = 0, rc2 = 0, Uo = 0;
    hipStream_t stream = nullptr;
    // tcp_md_steps runs the reciprocal part on stream2 beside combine_kernel:
    // forked from stream after the pair kernel, joined before the forces add
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // state
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    tcp::FixPos* upos = nullptr;
    // pair partials
    int nchunk = 1, jchunk = 1;
    int n3_rows = 0, n3_blocks = 0;  // Newton's-third-law kernel: block-row slots and blocks
    int2* blkmap = nullptr;
    int* n3_queue = nullptr;  // work queue of the persistent Newton kernel
    int n3_grid = 1;          // its resident blocks
    int n3_shape = 0;         // its block shape (0, or kFineN3Shape for small systems)
    size_t pf_cap = 0;    // doubles
    size_t part_cap = 0;  // doubles of part
    double *pf = nullptr, *pu = nullptr;
    // reciprocal part
    double *Q = nullptr, *C = nullptr, *gb = nullptr, *epart = nullptr, *sk = nullptr;
    double2* skpart = nullptr;
    double2* A = nullptr;
    double2* tw[3] = {nullptr, nullptr, nullptr};
    double *frx = nullptr, *fry = nullptr, *frz = nullptr;
    int nepart = 0;
    int nbrick[3] = {0, 0, 0};
    int *brick_of = nullptr, *brick_count = nullptr, *brick_start = nullptr;
    int *brick_cursor = nullptr, *sorted = nullptr;
    double* tiles = nullptr;
    // reductions and histograms
    double* part = nullptr;
    double* sc = nullptr;
    unsigned long long* rdf = nullptr;
    unsigned long long* vdf = nullptr;
    double* trace = nullptr;
    int trace_cap = 0;
    // external block (Force_direct_ext)
    int g_cap = 0;
    double *gx = nullptr, *gy = nullptr, *gz = nullptr;
    double *gfx = nullptr, *gfy = nullptr, *gfz = nullptr;
    tcp::FixPos* gupos = nullptr;
    double *gpf = nullptr, *gpu = nullptr;
    size_t gpf_cap = 0;
    // host arrays of tcp_set_state/get_state staged through pinned host memory,
    // moved in pipelined pieces and transposed to/from the SoA state on the device
    double* h_stage = nullptr;  // hipHostMalloc
    double* d_stage = nullptr;  // hipMalloc
    size_t stage_cap = 0;       // doubles in each
    // energy derivatives (allocated by the first tcp_energy_adjoint call)
    double *adj_part = nullptr, *adj_out = nullptr, *adj_q = nullptr, *adj_pot = nullptr;
    // timing
    hipEvent_t ev_pair[2][tcp::kEvPool];
    hipEvent_t ev_all[2];
    hipEvent_t ev_stage[tcp::kStageChunks];  // download chunks of the host staging
    bool events = false;
    double pair_ms = 0, steps_ms = 0;
    double ur_scale = 0.5;
    uint64_t counter = 0;
    std::vector<void*> allocs;
    std::string err;
};


namespace tcp {

MdArgs md_args(tcp_ctx* c);
RecipArgs recip_args(tcp_ctx* c);
PairArgs pair_args(tcp_ctx* c);
// j chunks of an ordered-pair launch: >= 2048 blocks, >= 128 j per chunk, <= 64 chunks
void choose_chunks(int n_i, int n_j, int* nchunk, int* jchunk);
// Parameter checks of tcp_create (no device use): 0 or a TCP_ERR_* code and *msg.
int validate_params(const tcp_params* p, int precision, std::string* msg);
// Sets the text tcp_last_error(NULL) returns on this thread; returns code.
int set_create_error(int code, const char* msg);

}  // namespace tcp

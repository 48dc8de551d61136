// C-ABI glue: error reporting, device information, the batched Thomas solver and the sweep entry points.
#include <hip/hip_runtime.h>

#include <cstdio>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "qz_device.hpp"
#include "qz_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

int qz_fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
int qz_fail_hip(hipError_t e, const std::string& where) {
    g_last_error = where + ": " + hipGetErrorString(e);
    return QZ_EHIP;
}

extern "C" const char* qz_last_error(void) { return g_last_error.c_str(); }

extern "C" int qz_device_info(char* arch, int arch_len, int* n_cu) {
    int dev = 0;
    QZ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    QZ_HIP(hipGetDeviceProperties(&prop, dev));
    if (arch && arch_len > 0) {
        std::strncpy(arch, prop.gcnArchName, static_cast<size_t>(arch_len) - 1);
        arch[arch_len - 1] = 0;
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    return QZ_OK;
}

namespace qz {
// Thomas, one thread per system (src/Tridiag.cpp:111-126), every value formed exactly as the reference forms it.
// Element i of system s sits at s * ss + i * es of a, b, c, d and x; the forward sweep keeps d' in x and c' in
// the caller's workspace (system s at s * M), which the backward sweep reads.
__global__ void thomas_batched_kernel(const Real* __restrict__ a, const Real* __restrict__ b, const Real* __restrict__ c,
                                      const Real* __restrict__ d, Real* __restrict__ x, Real* __restrict__ cpw, int M,
                                      int batch, long long ss, long long es) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= batch) return;
    const long long o = static_cast<long long>(s) * ss;
    const Real *A = a + o, *B = b + o, *C = c + o, *D = d + o;
    Real* X = x + o;
    Real* CP = cpw + static_cast<size_t>(s) * M;
    Real cp = C[0] / B[0];
    Real dp = D[0] / B[0];
    CP[0] = cp;
    X[0] = dp;
    for (int i = 1; i < M; ++i) {
        const long long e = i * es;
        const Real m = Real(1) / (B[e] - A[e] * cp);
        cp = C[e] * m;
        dp = (D[e] - A[e] * dp) * m;
        CP[i] = cp;
        X[e] = dp;
    }
    Real xn = X[(M - 1) * es];
    for (int i = M - 2; i >= 0; --i) {
        xn = X[i * es] - CP[i] * xn;
        X[i * es] = xn;
    }
}
}  // namespace qz

extern "C" size_t qz_tridiag_workspace_doubles(int M, int batch) {
    return (M < 1 || batch < 1) ? 0 : static_cast<size_t>(M) * static_cast<size_t>(batch);
}

extern "C" int qz_tridiag_thomas_batched(const double* a, const double* b, const double* c, const double* d, double* x,
                                         int M, int batch, long long sys_stride, long long elem_stride, double* work,
                                         void* stream) {
    if (M < 1 || batch < 0) return qz_fail(QZ_EINVAL, "qz_tridiag_thomas_batched: bad extent");
    if (batch == 0) return QZ_OK;
    if (!a || !b || !c || !d || !x || !work) return qz_fail(QZ_EINVAL, "qz_tridiag_thomas_batched: null argument");
    if (elem_stride < 1 || (batch > 1 && sys_stride < 1))
        return qz_fail(QZ_EINVAL, "qz_tridiag_thomas_batched: strides must be positive");
    // systems may interleave (element-major layouts), but no two elements of the batch may share an address:
    // (s, i) and (s + ds, i - di) collide when ds * sys_stride = di * elem_stride with di <= M - 1
    const long long span = (M - 1) * elem_stride + 1;
    if (batch > 1 && sys_stride < span)
        for (long long si = 1; si < batch && si * sys_stride < span; ++si)
            if ((si * sys_stride) % elem_stride == 0)
                return qz_fail(QZ_EINVAL, "qz_tridiag_thomas_batched: strides make two systems overlap");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int tb = 64;
    hipLaunchKernelGGL(qz::thomas_batched_kernel, dim3((batch + tb - 1) / tb), dim3(tb), 0, st, a, b, c, d, x, work, M,
                       batch, sys_stride, elem_stride);
    QZ_HIP(hipGetLastError());
    return QZ_OK;
}

extern "C" size_t qz_sweep_workspace_doubles(int M) { return qz_sweep_workspace_doubles_impl(M); }
extern "C" size_t qz_sweep_args_bytes(void) { return sizeof(qz_sweep_args); }

extern "C" int qz_device_tag(char* out, int n) {
    if (!out || n < 1) return qz_fail(QZ_EINVAL, "qz_device_tag: no buffer");
    int dev = 0;
    QZ_HIP(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    QZ_HIP(hipGetDeviceProperties(&pr, dev));
    std::snprintf(out, static_cast<size_t>(n), "%s|%dCU", pr.gcnArchName, pr.multiProcessorCount);
    return QZ_OK;
}

int qz_validate_sweep(const qz_sweep_args& a) {
    if (a.modes != 0 && a.modes != 1) return qz_fail(QZ_EINVAL, "sweep: modes must be 0 or 1");
    if (a.grid.N < 2 || a.grid.N > 2048) return qz_fail(QZ_ERANGE, "sweep: need 2 <= N <= 2048 radial cells");
    if (a.nslices < 1) return qz_fail(QZ_EINVAL, "sweep: nslices < 1");
    if (a.k_begin < 0 || a.k_end < 0 || a.k_end > a.nslices || (a.k_end > 0 && a.k_begin >= a.k_end) ||
        (a.k_end == 0 && a.k_begin >= a.nslices))
        return qz_fail(QZ_EINVAL, "sweep: slice range [k_begin, k_end) outside [0, nslices)");
    if (a.ab_order < 1 || a.ab_order > 5) return qz_fail(QZ_EINVAL, "sweep: ab_order must be 1..5");
    if (a.nspecies < 0 || a.nspecies > QZ_MAX_SPECIES) return qz_fail(QZ_ERANGE, "sweep: at most 2 mobile species");
    if (a.ncf != (a.modes ? 18 : 6) || a.ncb != (a.modes ? 15 : 5) || a.ncd < (a.modes ? 9 : 3))
        return qz_fail(QZ_EINVAL, "sweep: component counts do not match modes");
    if (!a.ws || !a.fld || !a.diag || !a.bsrc || !a.bg || !a.lost) return qz_fail(QZ_EINVAL, "sweep: null array");
    if (a.groups < 0) return qz_fail(QZ_EINVAL, "sweep: negative groups");
    if (a.load)
        for (int s = 0; s < a.nspecies; ++s) {
            const qz_plasma_rings& rg = a.rings[s];
            if (rg.ppc < 1 || rg.ntheta < 1 || rg.nring < 0 || rg.nring > a.grid.N * rg.ppc || !rg.cs ||
                rg.nring * rg.ntheta != a.sp[s].np)
                return qz_fail(QZ_EINVAL, "sweep: the ring load of a species does not match its particle count");
        }
    if (a.groups > 1 && !a.mws) return qz_fail(QZ_EINVAL, "sweep: a run of several workgroups needs its workspace mws");
    if (a.traj && a.groups > 1) return qz_fail(QZ_EINVAL, "sweep: the trajectory buffer needs a one-workgroup run");
    for (int s = 0; s < a.nspecies; ++s)
        if (a.sp[s].np < 0 || a.sp[s].np > a.sp[s].cap || !a.sp[s].hist) return qz_fail(QZ_EINVAL, "sweep: bad species");
    return QZ_OK;
}

// The kernel reads its argument blocks from device memory for its whole duration.  They are staged in one
// buffer per (device, stream), reused only by later calls on the same stream, which the stream orders after
// the kernels that read it.  (A stream-ordered allocation per call could be handed to another stream while a
// kernel still reads it.)  The lock keeps the copy and the launch of one call together.
namespace {
struct ArgsBuf {
    qz_sweep_args* p = nullptr;
    int cap = 0;
    int* map = nullptr;         // block -> (run, group) of a multi-workgroup launch
    int mapcap = 0;
    std::vector<int> hmap;      // host side of the map (alive while the copy is staged)
};
std::mutex g_args_mu;
std::map<std::pair<int, hipStream_t>, ArgsBuf> g_args;
}  // namespace

static int sweep_many(const qz_sweep_args* runs, int n, void* stream) {
    if (!runs || n < 1) return qz_fail(QZ_EINVAL, "sweep: no runs");
    for (int i = 0; i < n; ++i) {
        const int rc = qz_validate_sweep(runs[i]);
        if (rc) return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0;
    QZ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_args_mu);
    ArgsBuf& b = g_args[{dev, st}];
    if (b.cap < n) {   // grow: the old buffer may still be read by this stream's kernels
        if (b.p) {
            QZ_HIP(hipStreamSynchronize(st));
            QZ_HIP(hipFree(b.p));
            b.p = nullptr;
            b.cap = 0;
        }
        QZ_HIP(hipMalloc(reinterpret_cast<void**>(&b.p), sizeof(qz_sweep_args) * n));
        b.cap = n;
    }
    QZ_HIP(hipMemcpyAsync(b.p, runs, sizeof(qz_sweep_args) * n, hipMemcpyHostToDevice, st));
    bool multi = false;
    for (int i = 0; i < n; ++i) multi = multi || runs[i].groups > 1;
    if (!multi) return qz_sweep_launch(b.p, runs, n, st);
    // several workgroups per run: the block map, and every run's barrier counter and error flag at zero
    b.hmap = qz_sweep_block_map(runs, n);
    const int need = static_cast<int>(b.hmap.size());
    if (b.mapcap < need) {
        if (b.map) {
            QZ_HIP(hipStreamSynchronize(st));
            QZ_HIP(hipFree(b.map));
            b.map = nullptr;
            b.mapcap = 0;
        }
        QZ_HIP(hipMalloc(reinterpret_cast<void**>(&b.map), sizeof(int) * need));
        b.mapcap = need;
    }
    QZ_HIP(hipMemcpyAsync(b.map, b.hmap.data(), sizeof(int) * need, hipMemcpyHostToDevice, st));
    for (int i = 0; i < n; ++i) {
        if (!runs[i].mws) return qz_fail(QZ_EINVAL, "sweep: in a multi-workgroup launch every run needs mws");
        const int rc = qz_zero_async(runs[i].mws, 64 * sizeof(double), st);
        if (rc) return rc;
    }
    return qz_sweep_multi_launch(b.p, b.map, runs, n, st);
}

extern "C" int qz_sweep(const qz_sweep_args* args, void* stream) { return sweep_many(args, 1, stream); }
extern "C" int qz_sweep_release(void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    int dev = 0;
    QZ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_args_mu);
    auto it = g_args.find({dev, st});
    if (it == g_args.end()) return QZ_OK;
    QZ_HIP(hipStreamSynchronize(st));   // its kernels may still read the block
    if (it->second.p) QZ_HIP(hipFree(it->second.p));
    if (it->second.map) QZ_HIP(hipFree(it->second.map));
    g_args.erase(it);
    return QZ_OK;
}
extern "C" int qz_sweep_batched(const qz_sweep_args* runs, int nruns, void* stream) {
    return sweep_many(runs, nruns, stream);
}

// ---------------------------------------------------------------------------- the sweep without allocation
// For callers that cannot allocate device memory (code under HIP graph capture, serving loops): one run's sweep
// whose argument block and workspace live in this library's static device array `scratch`, so the call makes
// no allocation at all.  Every GPU holds its own copy of the array for the life of the process; its address on
// a device is looked up once, with that device current (hipGetSymbolAddress), and only that address is handed
// to kernels and copies.  Calls on one GPU share the array: two calls whose sweeps overlap in time race, and
// the workspace keeps what the last call left (a slice range continuing an earlier one reads its B+ from it).
namespace {
constexpr size_t kScratchArgBytes = 4096;                       // one qz_sweep_args block, rounded up
constexpr int kScratchMaxM = QZ_NOALLOC_MAX_NODES;              // grid nodes N + 1
constexpr size_t kScratchBytes = kScratchArgBytes + sizeof(double) * QZ_SWEEP_WS_ARRAYS * kScratchMaxM;
static_assert(sizeof(qz_sweep_args) <= kScratchArgBytes, "the argument block must fit its scratch slot");
}  // namespace

__device__ __align__(256) char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;   // device -> address of `scratch` on it

int scratch_on_current_device(char** out) {
    int dev = 0;
    QZ_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return qz_fail_hip(e, "qz_sweep_noalloc: hipGetSymbolAddress(scratch)");
        it = g_scratch_ptr.emplace(dev, static_cast<char*>(p)).first;
    }
    *out = it->second;
    return QZ_OK;
}
}  // namespace

namespace qz {
// the argument block goes to device memory as a kernel argument: a graph that captures the call keeps the values
// in its kernel node (a copy from host memory would read the caller's memory again at every replay)
__global__ void put_args_kernel(qz_sweep_args a, qz_sweep_args* __restrict__ dst) {
    constexpr int nw = static_cast<int>(sizeof(qz_sweep_args) / sizeof(unsigned int));
    const unsigned int* src = reinterpret_cast<const unsigned int*>(&a);
    unsigned int* out = reinterpret_cast<unsigned int*>(dst);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) out[i] = src[i];
}
}  // namespace qz

extern "C" int qz_sweep_noalloc(const qz_sweep_args* args, void* stream) {
    if (!args) return qz_fail(QZ_EINVAL, "qz_sweep_noalloc: null argument");
    const int M = args->grid.N + 1;
    if (M > kScratchMaxM)
        return qz_fail(QZ_ERANGE, "qz_sweep_noalloc: " + std::to_string(M) + " grid nodes exceed the scratch array's " +
                                      std::to_string(kScratchMaxM));
    char* base = nullptr;
    const int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    if (args->groups > 1) return qz_fail(QZ_EINVAL, "qz_sweep_noalloc: one workgroup per run (groups 0 or 1)");
    qz_sweep_args a = *args;
    a.ws = reinterpret_cast<double*>(base + kScratchArgBytes);   // the run's workspace: the scratch after the block
    const int v = qz_validate_sweep(a);
    if (v) return v;
    hipStream_t st = static_cast<hipStream_t>(stream);
    qz_sweep_args* dev_args = reinterpret_cast<qz_sweep_args*>(base);
    static_assert(sizeof(qz_sweep_args) % sizeof(unsigned int) == 0, "the block is copied in 4-byte words");
    hipLaunchKernelGGL(qz::put_args_kernel, dim3(1), dim3(64), 0, st, a, dev_args);
    QZ_HIP(hipGetLastError());
    return qz_sweep_launch(dev_args, &a, 1, st);
}

// Internal helpers of the quarz_amd library: error reporting and the launch functions of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/quarz_amd.h"

int qz_fail(int code, const std::string& msg);
int qz_fail_hip(hipError_t e, const std::string& where);

#define QZ_HIP(call)                                              \
    do {                                                          \
        hipError_t _e = (call);                                   \
        if (_e != hipSuccess) return qz_fail_hip(_e, #call);      \
    } while (0)

// arrays of M doubles in a run's sweep workspace, and the largest grid (nodes) of the sweep that allocates nothing
constexpr int QZ_SWEEP_WS_ARRAYS = 56;
constexpr int QZ_NOALLOC_MAX_NODES = 2049;
size_t qz_sweep_workspace_doubles_impl(int M);
// one workgroup per run, run b on workgroup b
int qz_sweep_launch(const qz_sweep_args* dev_runs, const qz_sweep_args* host_runs, int nruns, hipStream_t stream);
int qz_validate_sweep(const qz_sweep_args& a);   // the checks of qz_sweep, QZ_OK or a QZ_E* code
// the one-workgroup sweep holds a run of np plasma particles on N cells (m1: modes 0 and 1)
bool qz_sweep_fits_one_group(bool m1, int N, int np);
// G cooperating workgroups per run (qz_sweep_args.groups > 1); dev_map: block -> (run, group) pairs
int qz_sweep_multi_launch(const qz_sweep_args* dev_runs, const int* dev_map, const qz_sweep_args* host_runs, int nruns,
                          hipStream_t stream, int resident = 0);
// the block map of a multi-workgroup launch of these runs (2 ints per block)
std::vector<int> qz_sweep_block_map(const qz_sweep_args* runs, int nruns);
// particles per thread (2, 4 or 8) the multi-workgroup kernel needs for np particles on `groups` workgroups; one
// launch takes the largest of its runs', and workgroup g of every run owns particles [g P, (g + 1) P),
// P = 512 x that value, so runs of one launch are best given workgroup counts of one value
int qz_sweep_multi_pt(int np, int groups);
namespace qz {
long sweep_lds_needed(bool m1, int N, int pt, int L);   // LDS bytes of the one-workgroup kernel
long sweep_multi_lds_needed(bool m1, int N, int pt);    // LDS bytes of the multi-workgroup kernel
}  // namespace qz
// One launch of a batch with runs of one and of several workgroups (qz_sweep_mixed.hip): `slots` blocks for the
// multi-workgroup runs (dev_map: 2 ints per slot; run -1: a slot that works the queue, -2: one that ends), then
// `workers` blocks; every block without a run takes the one-workgroup runs from the device counter `queue`, in
// order, each swept with its own particles per thread.  Every multi-workgroup run must need 2 particles per
// thread (qz_sweep_multi_pt).  nmulti = 0, slots = 0: a queue of one-workgroup runs only.
int qz_sweep_mixed_launch(const qz_sweep_args* dev_single, const qz_sweep_args* host_single, int nsingle,
                          unsigned int* queue, const qz_sweep_args* dev_multi, const int* dev_map,
                          const qz_sweep_args* host_multi, int nmulti, int slots, int workers, hipStream_t stream);
int qz_beam_count_live(int np, const double* w, unsigned long long* count, hipStream_t st);
// qz_beam_deposit_beams into the rows [box->k0, box->k0 + box->nloc) at bsrc, overwriting them (every node of
// the window written, nodes without particles zero) instead of adding to what they hold: no zero fill first
int qz_beam_deposit_window(const qz_radial_grid* grid, const qz_beam_grid* box, int modes, int nbeams,
                           const double* q, const int* np, const qz_beam_soa* p, double* bsrc, int nc, void* work,
                           const int* const* pstep, int marker, unsigned long long* live_count, hipStream_t st);
// zero `bytes` (a multiple of 4) at p with a kernel: unlike a memset node it is safe inside a captured graph
int qz_zero_async(void* p, size_t bytes, hipStream_t st);
// out[r][j] = s0 a[K-1-r][j][c0] + s1 a[K-1-r][j][c1] (s1 = 0: one term), a of K x M x nc doubles, on the device
int qz_field_component(const double* a, int nc, int K, int M, int c0, double s0, int c1, double s1, double* out,
                       hipStream_t st);

// Device-resident model copies, per-handle workspace and batch descriptors.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "sus2_device.hpp"
#include "sus2_model.hpp"

// Model shapes with compiled kernel instances: (L, K, radial basis size B,
// warp sharing 0 pair / 1 L / 2 K / 3 LK).  Adding a shape is one entry here.
#define SUS2_FOR_EACH_SHAPE(X) X(1, 2, 3, 3) X(3, 4, 3, 3) X(3, 4, 10, 2) X(4, 6, 10, 2) X(5, 6, 10, 2)

namespace sus2 {

// A caller-provided region of device memory that buffers are carved from in
// order (no allocation): while a thread has an arena installed (ArenaScope),
// every DevBuf::ensure of that thread takes its bytes from it.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};
Arena* current_arena();
// While a thread captures a HIP graph, buffer growth (hipMalloc / hipFree)
// and host waits are errors: every buffer is sized before the capture.
bool capturing();
struct CaptureScope {
  CaptureScope();
  ~CaptureScope();
  CaptureScope(const CaptureScope&) = delete;
  CaptureScope& operator=(const CaptureScope&) = delete;
};
struct ArenaScope {
  explicit ArenaScope(Arena* a);
  ~ArenaScope();
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;
  Arena* prev;
};

// Growable device allocation (never shrinks; freed with its owner), or a
// region carved from the installed arena (not freed).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void ensure(size_t n);
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
    std::swap(owned, o.owned);
  }
  template <class U> U* get() const { return static_cast<U*>(p); }
};

struct Batch {
  int nframes = 0, natoms = 0;
  const int64_t* frame_off = nullptr;
  const int32_t* types = nullptr;
  const double* pos = nullptr;
  const double* cell = nullptr;
  const int32_t* periodic = nullptr;
};

struct EfsOut {
  double *energy = nullptr, *force = nullptr, *virial = nullptr, *site = nullptr;
};

// Optional per-kernel timing: an event pair around every launch on the
// launch stream, accumulated per kernel name when the events complete.
struct Profiler {
  bool on = false;
  struct Pending { int id; hipEvent_t a, b; };
  std::vector<std::string> names;
  std::vector<double> total_ms;
  std::vector<long> launches;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  ~Profiler();
  // returns -1 when off, else the index of a fresh start/stop event pair
  int begin(const std::string& name);
  hipEvent_t start_event(int idx) const { return pending[idx].a; }
  hipEvent_t stop_event(int idx) const { return pending[idx].b; }
  void collect_ready();  // accumulate completed pairs (non-blocking)
  void collect_all();    // wait for and accumulate every pending pair
  void reset();
};

struct Workspace {
  Profiler prof;
  DevBuf nbr_cnt, nbr_off, scan_tmp, rec, h, abar, gbar, vg, site, fw;
  // per-edge backward buffers: gedge_in, uedge hold dE/dd of each directed
  // edge and its gate-adjoint message at the edge's slot in its target's
  // incoming list (in_pos), so that gathers read contiguous ranges; in_edges
  // the sorted incoming-edge lists of the atomic placement path
  DevBuf gedge_in, uedge, in_edges, in_pos, in_cur;
  // fixed-capacity per-atom neighbour slots of the single-pass list build
  DevBuf nbr_slots;
  DevBuf nbr_oslot;  // bin-offset index of each slot's hit (ascending per atom)
  int nbr_cap = 128;
  // cell-list neighbour build
  DevBuf grid, cell_of, wrap, cell_cnt, cell_start, cell_cur, cell_atoms;
  DevBuf cell_pack;   // [N] bin-ordered atoms: home-cell position, index, species
  DevBuf atom_home;   // [N] home-cell position and bin id per atom
  DevBuf nbr_otab;    // [N][125] uint16 hits before each bin offset
  DevBuf atom_frame;  // [N] frame index of every atom (written by the binning kernels)
  // parameter gradient: gate tangent h_dot, gate node tangents, type-free gate
  // moments and tangents, per-edge gate messages, gathered h_bar, per-workgroup
  // FP64 partials (coefficients; gate weights W)
  DevBuf hdot, vgd, gmom, gmomd, msg, hbar, gpart, wpart, gfold;
  DevBuf rec_alt;  // FP32 copy of FP64 edge records (mixed-precision objective)
  // training objective: EFS outputs, loss adjoints and per-frame losses (FP64)
  DevBuf lE, lF, lV, lew, lfw, lsw, lloss, lsite, lpen;
  long nedges = 0;
  // edge capacity (edges, not bytes) fixed by the last host-sized build: the
  // per-edge buffers of either precision are sized from it, so that a build
  // sized from earlier calls neither creeps past their slack nor overflows
  // after a change of precision (FP64 records are twice as wide)
  long edge_cap = 0;
  long max_frame_atoms = 0;  // upper bound on the atoms of one frame of the current batch
  DevBuf fin_part;           // per-chunk partial frame sums of the frame reduction
  DevBuf fin_done;           // per-frame count of finished chunks (zero between calls)
  int64_t* host_total = nullptr;  // pinned [4]
  // Sizing of the neighbour build without a host round trip: a build may take
  // its edge capacity from earlier calls (nbr_known) when the caller verifies
  // the device's overflow report afterwards (neighbors_verify); nbr_pending
  // marks such a build whose report has not been read yet.
  bool nbr_optimistic = false, nbr_known = false, nbr_pending = false;
  // Capture-safe mode (HIP graph capture): no host waits inside a call.  The
  // neighbour build takes its edge capacity from earlier calls and reports an
  // overflow only through the sticky device flag overflow_flag (the lists are
  // then empty and the outputs invalid); the largest frame of the batch is
  // known_max_frame; a training objective goes to a device word only.
  bool capture = false;
  int* overflow_flag = nullptr;
  long known_max_frame = -1;
  // Count report of such a build, written by the scan kernel into coherent
  // pinned host memory and polled by the host: two 64-bit words, each tagged
  // with the build's sequence number in its upper half (word 0: largest count;
  // word 1: overflow flag in bit 31, total in bits 0-30, saturated); word 2
  // holds the largest frame of a batch, words 4-5 the low and high halves of
  // a training objective, tagged the same way.  Kept until the process ends
  // (a device-written host page is never freed).
  int64_t* host_report = nullptr;
  int64_t report_seq = 0, scan_seq = 0;  // last tag issued; tag of the pending scan report
  ~Workspace();
  void ensure_atoms(int n);
};

// exclusive scan cnt[0..n) -> off[0..n] on the device (total[0] = sum, total[1] = max)
// With cap_lim / edge_lim: when the largest count exceeds cap_lim or the sum
// exceeds edge_lim, total[2] = 1 and every count and offset is written as 0
// (an empty list for every atom); otherwise total[2] = 0.
// With report: the largest count and the total with the overflow flag are
// also written to report[0..1], each word tagged with seq in its upper half.
// With sticky: set to 1 (and never cleared here) when the lists overflowed.
void device_scan(int* cnt, int* off, int n, int64_t* total, hipStream_t st, int cap_lim = INT32_MAX,
                 long edge_lim = INT64_MAX, int64_t* report = nullptr, int64_t seq = 0, int* sticky = nullptr);
// exclusive scan of ws.nbr_cnt[0..n) into ws.nbr_off[0..n]; returns the total
// (and the largest count in *maxcnt)
long exclusive_scan(Workspace& ws, int n, hipStream_t st, long* maxcnt = nullptr);
// Device address of the host report's word k (for the kernels that write it).
int64_t* report_word(Workspace& ws, int k);
// Polls words [k, k + n) of the host report until each carries tag seq in its
// upper half and returns them in w; spins for up to 50 ms, then waits for the
// stream.  Throws when the stream fails or drains without the report.
void poll_report(Workspace& ws, int k, int n, int64_t seq, hipStream_t st, const char* what, uint64_t* w);

// One precision's device view of a model.
template <class T>
struct DevModelT {
  DevBuf theta, rows;  // coefficient vector; product-graph rows (gate forward, outer forward, gate reverse, outer reverse)
  DModel<T> view{};
};

// Launch shape of the per-atom EFS kernels (forward passes, edge backwards,
// gathers): centres per edge-backward group (1, 2 or 4) and the grid, one
// workgroup per atom / group (grid_mult 0) or grid_mult times the resident
// workgroups striding over them (persistent).  Set by the autotuner or by the
// caller; the default is the untuned shape.
struct EfsTune {
  int eb_atoms = 2;
  int grid_mult = 0;
  // gate-adjoint messages of the batch path of Effect models: 0 written per
  // edge by the outer edge backward, 1 recomputed by the outer gather from the
  // records and the neighbours' moment adjoints (no per-edge message buffer),
  // 2 written per edge by a messages-only pass after a forces-only one; -1
  // automatic: 2 from L = 4 on (the combined pass is register-bound there),
  // else 0
  int msg_recompute = -1;
};

struct DevModelSet {
  const Model* host = nullptr;
  EfsTune tune[2];  // [0] FP64, [1] FP32 kernels
  template <class T> const EfsTune& tune_of() const { return tune[sizeof(T) == 4 ? 1 : 0]; }
  DevBuf ints;  // all integer schedule arrays
  DevModelT<double> d;
  DevModelT<float> f;
  template <class T> const DModel<T>& get() const;
  void upload(const Model& m);
  void upload_theta(const Model& m);
};
template <> inline const DModel<double>& DevModelSet::get<double>() const { return d.view; }
template <> inline const DModel<float>& DevModelSet::get<float>() const { return f.view; }

void efs_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const EfsOut& out,
                  hipStream_t st);
// grad (+= when add, else =) scale * parameter gradient of the weighted EFS
// combination; reuse_neighbors skips the neighbour build when the batch's list
// of this precision is current (built by the preceding EFS call).
void grad_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* ene_w,
                   const double* frc_w, const double* str_w, double* grad, double scale, bool add,
                   bool reuse_neighbors, hipStream_t st);
// Objective options of a handle (the NonLinearRegression members the
// trainer sets; sus2_objective in the C ABI).  The defaults are those of
// `mlp-sus2 train` for Static and Effect models.
struct ObjectiveOpts {
  double rel_forces = 0.0;     // wgt_rel_forces
  int norm_by_forces = 0;      // norm_by_forces
  int force_loss = 0;          // 0 L2, 1 log-cosh
  double log_cosh_scale = 2.0;
  double std_weight = 0.0, stdd_weight = 0.0;
  double constr_weight = 1e-9;  // legacy radial norm / orthogonality penalty
  bool std_terms() const { return std_weight != 0.0 || stdd_weight != 0.0; }
  void validate() const;
};
// The penalty constr (sum_mu (|c_mu|^2 - 1)^2 + sum_{mu<nu} (c_mu . c_nu)^2)
// over the outer radial rows of theta, and (grad != nullptr) its gradient
// added to grad (host FP64; MLMTPR::AddPenaltyGrad, dev_src/mtpr.cpp:14050).
double radial_penalty(const Model& m, const double* theta, double constr, double* grad);
// Training objective (1/nframes) sum_f L_f + penalty of the batch with its
// parameter gradient (grad overwritten); labels e0/f0/s0 on the device, loss
// to host.  loss_dev (optional): the objective is also written to this device
// word; efs_out (optional): the EFS outputs of the batch go to these buffers.
// With loss_host null or in capture-safe mode nothing is read back.
void loss_grad_dispatch(const DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* e0,
                        const double* f0, const double* s0, int mode, double wE, double wF, double wS,
                        const ObjectiveOpts& obj, double* loss_host, double* grad, hipStream_t st,
                        double* loss_dev = nullptr, const EfsOut* efs_out = nullptr);

// Linear readout solve (species constants and per-element readout weights)
// of the training objective (its norm_by_forces scaling and penalty; the
// other options are not part of the reference's normal equations); writes the
// solution into host.theta and the device copies, returns the objective at
// the solution.
double train_linear(Model& host, DevModelSet& dm, Workspace& ws, int precision, const Batch& b, const double* e0,
                    const double* f0, const double* s0, int mode, double wE, double wF, double wS, double ridge,
                    const ObjectiveOpts& obj, hipStream_t st);

}  // namespace sus2

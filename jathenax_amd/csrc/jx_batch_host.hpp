// Host side of the batched engine: instance slots in one HBM pool, the tile list, the
// per-round launch sequence and its hipGraph.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../../include/jathenax_amd.h"
#include "jx_batch.hpp"

namespace jx {

struct BatchBase {
  virtual ~BatchBase() {}
  std::string err;
  virtual int add(const jx_run_params* p, const void* U0, const jx_forcing* f, int rows,
                  int* id, const jx_loop_state* start = nullptr) = 0;
  virtual int checkpoint(int id, void* U_out, jx_loop_state* state) = 0;
  virtual int remove(int id) = 0;
  virtual int run(int rounds, int* active) = 0;
  virtual int launch(int rounds) = 0;
  virtual int wait(int* active) = 0;
  virtual int poll(int* done, double* time, int* cycle, int cap) = 0;
  virtual int poll_launched(int* done, double* time, int* cycle, int cap, int* active) = 0;
  virtual int result(int id, void* U_out, double* time, double* dt, int* cycle) = 0;
  virtual int state_ptr(int id, void** p) = 0;
  virtual int set_graph(int on, int rounds) = 0;
  virtual int set_fill(int fill) = 0;
  virtual long long pool_free() const = 0;
  virtual long long footprint(const jx_run_params* p, int forced) const = 0;
};

inline bool batch_pow2(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return false;
  int e = 0;
  return std::frexp(x, &e) == 0.5;
}

template <typename T> class Batch : public BatchBase {
 public:
  Batch(int device, int max_inst, long long pool_bytes)
      : device_(device), max_inst_(max_inst), pool_bytes_(pool_bytes) {}

  int init() {
    if (hipSetDevice(device_) != hipSuccess) return fail("hipSetDevice failed");
    if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess)
      return fail("stream creation failed");
    hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, device_);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_stage_kernel<T>,
                                                     kBatchBlock, smem()) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    grid_ = std::max(1, cus_) * per_cu;
    nwaves_ = grid_ * (kBatchBlock / 64);
    if (hipMalloc(&d_inst_, sizeof(BatchInst<T>) * max_inst_) != hipSuccess ||
        hipMalloc(&d_ctrl_, sizeof(BatchCtrl) * max_inst_) != hipSuccess ||
        hipMalloc(&d_meta_, sizeof(int) * 8) != hipSuccess ||
        hipMalloc(&d_cnt_, sizeof(unsigned) * 8) != hipSuccess ||
        hipMalloc(&d_lists_, sizeof(int) * WL_TOTAL * list_stride(max_inst_)) != hipSuccess ||
        hipMalloc(&d_flags_, sizeof(int) * max_inst_) != hipSuccess ||
        hipMalloc(&d_pool_, pool_bytes_) != hipSuccess ||
        (Fast<T>::value &&
         (hipMalloc(&d_groups_, sizeof(int) * kGroupTable * max_inst_) != hipSuccess ||
          hipMalloc(&d_gscratch_, sizeof(int) * mode_group_ints(kMaxGroupModes)) !=
              hipSuccess)) ||
        hipHostMalloc(&h_ctrl_, sizeof(BatchCtrl) * max_inst_, hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(&h_meta_, sizeof(int) * 8, hipHostMallocDefault) != hipSuccess)
      return fail("batch allocation failed (pool too large for the device?)");
    for (int k = 0; k < 2; ++k)
      if (hipHostMalloc(&h_snap_[k], sizeof(BatchCtrl) * max_inst_, hipHostMallocDefault) !=
              hipSuccess ||
          hipHostMalloc(&h_snap_meta_[k], sizeof(int) * 8, hipHostMallocDefault) != hipSuccess ||
          hipEventCreateWithFlags(&snap_ev_[k], hipEventDisableTiming) != hipSuccess)
        return fail("batch snapshot allocation failed");
    hipMemsetAsync(d_ctrl_, 0, sizeof(BatchCtrl) * max_inst_, stream_);
    hipMemsetAsync(d_meta_, 0, sizeof(int) * 8, stream_);
    hipMemsetAsync(d_cnt_, 0, sizeof(unsigned) * 8, stream_);
    hipMemsetAsync(d_lists_, 0, sizeof(int) * WL_TOTAL * list_stride(max_inst_), stream_);
    hipMemsetAsync(d_flags_, 0, sizeof(int) * max_inst_, stream_);
    free_[0] = (size_t)pool_bytes_;
    h_inst_.resize(max_inst_);
    slots_.assign(max_inst_, Slot{});
    done_seen_.assign(max_inst_, 0);
    if (hipStreamSynchronize(stream_) != hipSuccess) return fail("init sync failed");
    return JX_OK;
  }

  ~Batch() override {
    if (stream_) hipStreamSynchronize(stream_);
    drop_graph();
    for (void* p : {(void*)d_inst_, (void*)d_ctrl_, (void*)d_meta_, (void*)d_cnt_,
                    (void*)d_pool_, (void*)d_lists_, (void*)d_flags_, (void*)d_groups_,
                    (void*)d_gscratch_})
      if (p) hipFree(p);
    if (h_ctrl_) hipHostFree(h_ctrl_);
    if (h_meta_) hipHostFree(h_meta_);
    for (int k = 0; k < 2; ++k) {
      if (h_snap_[k]) hipHostFree(h_snap_[k]);
      if (h_snap_meta_[k]) hipHostFree(h_snap_meta_[k]);
      if (snap_ev_[k]) hipEventDestroy(snap_ev_[k]);
    }
    if (stream_) hipStreamDestroy(stream_);
  }

  long long footprint(const jx_run_params* p, int forced) const override {
    const size_t n = (size_t)p->nx * p->ny;
    size_t b = 4 * align(n * p->nvar * sizeof(T));
    if (forced) b += align(n * 2 * sizeof(T)) + align(group_bytes(p->nx, p->ny));
    return (long long)b;
  }

  long long pool_free() const override {
    long long f = 0;
    for (auto& kv : free_) f += (long long)kv.second;
    return f;
  }

  int add(const jx_run_params* p, const void* U0, const jx_forcing* f, int rows,
          int* id, const jx_loop_state* start = nullptr) override {
    if (!p || !U0 || !id) return fail("add: null argument");
    if (start && (!std::isfinite(start->time) || start->cycle < 0 || start->update < -1 ||
                  !(start->prev_dt > 0.0)))
      return fail("add: bad loop state");
    if (p->dtype != (Fast<T>::value ? 1 : 0)) return fail("add: dtype does not match batch");
    if (p->nx < 1 || p->ny < 1) return fail("add: grid must be non-empty");
    if (p->nvar < 3 || p->nvar > kBatchNV) return fail("add: batched instances have nvar 3 or 4");
    if (p->reconstruct < 0 || p->reconstruct > 4 || p->rsolver < 0 || p->rsolver > 2 ||
        p->integrator < 2 || p->integrator > 4)
      return fail("add: unknown operator id", JX_EKEY);
    if (p->rk_rows && !rk_rows_are_builtin(p->rk_rows, p->integrator))
      return fail("add: the batched engine runs the built-in rk2/rk3/rk4 rows; this RK table "
                  "differs from them");
    for (double v : {p->viscosity, p->smagorinsky_cs, p->drag_rate, p->dedt})
      if (!std::isfinite(v) || v < 0.0) return fail("add: coefficients must be finite and >= 0");
    if (f && (!f->coefficients || !f->cos_x || !f->sin_x || !f->cos_y || !f->sin_y ||
              f->n_modes < 1 || f->n_updates < 1 || !(p->dt_update > 0.0)))
      return fail("add: incomplete forcing");
    int s = -1;
    for (int k = 0; k < max_inst_; ++k)
      if (!slots_[k].live) { s = k; break; }
    if (s < 0) return fail("add: no free instance slot");
    const size_t n = (size_t)p->nx * p->ny;
    const size_t sb = align(n * p->nvar * sizeof(T)), rb = f ? align(n * 2 * sizeof(T)) : 0;
    const size_t gb = f ? align(group_bytes(p->nx, p->ny)) : 0;
    const size_t need = 4 * sb + rb + gb;
    long long off = alloc(need);
    if (off < 0) return fail("add: pool exhausted");
    Slot& sl = slots_[s];
    sl = Slot{};
    sl.live = true;
    sl.since = launches_;
    done_seen_[s] = 0;
    sl.off = (size_t)off;
    sl.bytes = need;
    sl.p = *p;
    char* base = static_cast<char*>(d_pool_) + off;
    BatchInst<T> I{};
    I.A = reinterpret_cast<T*>(base);
    I.B = reinterpret_cast<T*>(base + sb);
    I.C = reinterpret_cast<T*>(base + 2 * sb);
    I.S = reinterpret_cast<T*>(base + 3 * sb);
    I.raw = f ? reinterpret_cast<T*>(base + 4 * sb) : nullptr;
    I.part = f ? reinterpret_cast<double*>(base + 4 * sb + rb) : nullptr;
    I.reg_stride = (long long)(sb / sizeof(T));
    I.nstrips = (p->nx + kTileCols - 1) / kTileCols;
    I.ngroups = (p->ny + kGroupRows - 1) / kGroupRows;
    // Band heights are multiples of the 8-row group; 0 asks for the default cap.
    I.rows_cap = std::max(kGroupRows, ((rows > 0 ? rows : 32) / kGroupRows) * kGroupRows);
    if (f) {
      I.coef = static_cast<const T*>(f->coefficients);
      I.cx = static_cast<const T*>(f->cos_x);
      I.sx = static_cast<const T*>(f->sin_x);
      I.cy = static_cast<const T*>(f->cos_y);
      I.sy = static_cast<const T*>(f->sin_y);
      if (Fast<T>::value && f->n_modes <= kMaxGroupModes) {
        // Mode groups of the x basis (render_grouped_kernel): found in the shared scratch,
        // then the part the render reads (G, gstart, order, sign) is kept per slot.  Adds
        // are ordered on the batch stream, so one scratch serves them all.
        const int M = f->n_modes;
        if (M > 1)
          hipLaunchKernelGGL(mode_relation_kernel<T>, dim3(M, M), dim3(kAuxBlock), 0, stream_,
                             I.cx, I.sx, p->nx, M, d_gscratch_);
        hipLaunchKernelGGL(mode_groups_kernel, dim3(1), dim3(1), 0, stream_, M, d_gscratch_);
        int* tab = d_groups_ + (size_t)s * kGroupTable;
        if (hipMemcpyAsync(tab, d_gscratch_, sizeof(int) * (2 + 3 * (size_t)M),
                           hipMemcpyDeviceToDevice, stream_) != hipSuccess)
          return fail("add: mode group copy failed");
        I.grp = tab;
      }
      I.n_updates = f->n_updates;
      I.n_modes = f->n_modes;
    }
    I.nx = p->nx;
    I.ny = p->ny;
    I.nv = p->nvar;
    I.recon = p->reconstruct;
    I.rs = p->rsolver;
    I.nstages = p->integrator;
    I.visc_on = p->viscosity > 0.0 || p->smagorinsky_cs > 0.0;
    I.smag_on = p->smagorinsky_cs > 0.0;
    I.drag_on = p->drag_rate > 0.0;
    I.forced = f != nullptr;
    I.dx = T(p->dx);
    I.dy = T(p->dy);
    I.cs = T(p->cs);
    I.cs2 = T(p->cs * p->cs);
    I.inv_cs = T(1.0 / p->cs);
    I.dfloor = T(p->dfloor);
    I.eps = T(std::max(1e-42, (double)std::numeric_limits<T>::min()));
    I.four_dx = T(4.0 * p->dx);
    I.four_dy = T(4.0 * p->dy);
    I.visc = T(p->viscosity);
    I.smag_c = T(p->smagorinsky_cs * p->smagorinsky_cs * p->dx * p->dy);
    I.drag = T(p->drag_rate);
    I.inv_dx = T(1.0 / p->dx);
    I.inv_dy = T(1.0 / p->dy);
    I.inv_four_dx = T(1.0 / (4.0 * p->dx));
    I.inv_four_dy = T(1.0 / (4.0 * p->dy));
    LoopArgs& L = I.loop;
    L.tlim = p->tlim;
    L.cfl = p->cfl;
    L.viscosity = p->viscosity;
    L.drag_rate = p->drag_rate;
    L.h = std::min(p->dx, p->dy);
    L.dt_update = p->dt_update;
    L.tdriv_start = p->tdriv_start;
    L.dedt = f ? p->dedt : 0.0;
    L.nlim = p->nlim;
    L.smag_on = I.smag_on;
    L.has_forcing = I.forced;
    L.n_updates = I.n_updates;
    h_inst_[s] = I;
    if (hipMemcpyAsync(I.A, U0, n * p->nvar * sizeof(T), hipMemcpyDeviceToDevice, stream_) !=
        hipSuccess)
      return fail("add: state copy failed");
    if (I.raw) hipMemsetAsync(I.raw, 0, n * 2 * sizeof(T), stream_);
    int rc = upload_tables();
    if (rc) return rc;
    BatchTabs<T> tb = tabs();
    hipLaunchKernelGGL(batch_init_kernel<T>, dim3(1), dim3(1), 0, stream_, tb, s);
    if (start) {
      const int upd = f ? start->update : -1;
      hipLaunchKernelGGL(batch_resume_kernel<T>, dim3(1), dim3(1), 0, stream_, tb, s,
                         start->time, start->prev_dt, start->cycle, upd);
      if (upd >= 0)
        hipLaunchKernelGGL(batch_render_slot_kernel<T>, dim3(grid_), dim3(kBatchBlock), 0,
                           stream_, tb, s);
    }
    hipLaunchKernelGGL(batch_timestep_kernel<T>, dim3(grid_), dim3(kBatchBlock), 0, stream_, tb,
                       s);
    if (hipGetLastError() != hipSuccess) return fail("add: init launch failed");
    *id = s;
    return JX_OK;
  }

  int remove(int id) override {
    if (id < 0 || id >= max_inst_ || !slots_[id].live) return fail("remove: no such instance");
    hipLaunchKernelGGL(batch_retire_kernel<T>, dim3(1), dim3(1), 0, stream_, tabs(), id);
    release(slots_[id].off, slots_[id].bytes);
    slots_[id] = Slot{};
    return upload_tables();
  }

  int run(int rounds, int* active) override {
    const int rc = launch(rounds);
    return rc ? rc : wait(active);
  }

  int launch(int rounds) override {
    if (rounds < 0) return fail("run: rounds must be >= 0");
    const int kind = round_kind();
    if (use_graph_) {
      if (!graph_exec_[kind]) {
        int rc = capture(kind);
        if (rc) return rc;
      }
      const int launches = (rounds + graph_rounds_ - 1) / graph_rounds_;
      for (int k = 0; k < launches; ++k)
        if (hipGraphLaunch(graph_exec_[kind], stream_) != hipSuccess)
          return fail("graph launch failed");
    } else {
      for (int k = 0; k < rounds; ++k) launch_round(kind);
    }
    if (hipGetLastError() != hipSuccess) return fail("round launch failed");
    // Snapshot of the loop state as of the end of these rounds (poll_launched).  With two
    // launches already unpolled, the older one's snapshot is dropped.
    if (snap_count_ == 2) {
      snap_tail_ ^= 1;
      snap_count_ = 1;
    }
    const int k = snap_head_;
    snap_seq_[k] = launches_++;
    if (hipMemcpyAsync(h_snap_[k], d_ctrl_, sizeof(BatchCtrl) * max_inst_,
                       hipMemcpyDeviceToHost, stream_) != hipSuccess ||
        hipMemcpyAsync(h_snap_meta_[k], d_meta_, sizeof(int) * 8, hipMemcpyDeviceToHost,
                       stream_) != hipSuccess ||
        hipEventRecord(snap_ev_[k], stream_) != hipSuccess)
      return fail("launch: snapshot failed");
    snap_head_ ^= 1;
    ++snap_count_;
    return JX_OK;
  }

  int wait(int* active) override {
    if (hipMemcpyAsync(h_meta_, d_meta_, sizeof(int) * 8, hipMemcpyDeviceToHost, stream_) !=
            hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
      return fail("run: synchronisation failed");
    drop_snapshots();   // everything launched has been waited for
    if (active) *active = h_meta_[2];
    return JX_OK;
  }

  int poll(int* done, double* time, int* cycle, int cap) override {
    if (hipMemcpyAsync(h_ctrl_, d_ctrl_, sizeof(BatchCtrl) * max_inst_, hipMemcpyDeviceToHost,
                       stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
      return fail("poll failed");
    drop_snapshots();
    read_ctrl(h_ctrl_, done, time, cycle, cap);
    return JX_OK;
  }

  // Empties the snapshot ring: the next launch's snapshot is the next one polled.
  void drop_snapshots() {
    snap_tail_ = snap_head_;
    snap_count_ = 0;
  }

  // The loop state as of the end of the oldest launch not yet polled: waits for that
  // launch's snapshot only, so later launches keep the GPU busy meanwhile.
  int poll_launched(int* done, double* time, int* cycle, int cap, int* active) override {
    if (!snap_count_) return fail("poll_launched: no launch is waiting to be polled");
    const int k = snap_tail_;
    if (hipEventSynchronize(snap_ev_[k]) != hipSuccess)
      return fail("poll_launched: synchronisation failed");
    read_ctrl(h_snap_[k], done, time, cycle, cap, snap_seq_[k]);
    if (active) *active = h_snap_meta_[k][2];
    snap_tail_ ^= 1;
    --snap_count_;
    return JX_OK;
  }

  // `seq`: the launch a snapshot was taken at (the current state: every launch so far).  A
  // slot whose instance was added after that launch shows its previous occupant there: it
  // is reported running, at time NaN and cycle -1, and its done flag is not recorded.
  void read_ctrl(const BatchCtrl* h, int* done, double* time, int* cycle, int cap,
                 long long seq = std::numeric_limits<long long>::max()) {
    for (int k = 0; k < std::min(cap, max_inst_); ++k) {
      const bool live = slots_[k].live;
      if (live && seq < slots_[k].since) {
        if (done) done[k] = 0;
        if (time) time[k] = std::numeric_limits<double>::quiet_NaN();
        if (cycle) cycle[k] = -1;
        continue;
      }
      if (live) done_seen_[k] = h[k].c.done ? 1 : 0;
      if (done) done[k] = live ? (h[k].c.done ? 1 : 0) : -1;
      if (time) time[k] = h[k].c.time;
      if (cycle) cycle[k] = h[k].c.cycle;
    }
  }

  int result(int id, void* U_out, double* time, double* dt, int* cycle) override {
    if (id < 0 || id >= max_inst_ || !slots_[id].live) return fail("result: no such instance");
    const jx_run_params& p = slots_[id].p;
    if (U_out &&
        hipMemcpyAsync(U_out, h_inst_[id].A, (size_t)p.nx * p.ny * p.nvar * sizeof(T),
                       hipMemcpyDeviceToDevice, stream_) != hipSuccess)
      return fail("result: copy failed");
    if (hipMemcpyAsync(&h_ctrl_[id], &d_ctrl_[id], sizeof(BatchCtrl), hipMemcpyDeviceToHost,
                       stream_) != hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
      return fail("result: synchronisation failed");
    if (time) *time = h_ctrl_[id].c.time;
    if (dt) *dt = h_ctrl_[id].c.prev_dt;
    if (cycle) *cycle = h_ctrl_[id].c.cycle;
    return JX_OK;
  }

  // State and loop state between rounds.  An instance whose final stage has run but whose
  // step the next control round would commit is reported committed (time + dt, cycle + 1);
  // an instance in mid-step is reported at the start of that step, whose state A still
  // holds (only the round of a step's final stage writes A: the final stage itself or, for
  // a forced Smagorinsky instance, the timestep pass that kicks its output).
  int checkpoint(int id, void* U_out, jx_loop_state* st) override {
    if (!st) return fail("checkpoint: null state");
    double t = 0.0, dt = 0.0;
    int cyc = 0;
    int rc = result(id, U_out, &t, &dt, &cyc);
    if (rc) return rc;
    const BatchCtrl& C = h_ctrl_[id];
    *st = jx_loop_state{C.c.time, C.c.prev_dt, C.c.cycle, C.c.update};
    if (C.c.pending && C.active && C.stage == h_inst_[id].nstages - 1 && !C.c.done) {
      st->time = C.c.time + C.c.dt;
      st->prev_dt = C.c.dt;
      st->cycle = C.c.cycle + 1;
    }
    return JX_OK;
  }

  int state_ptr(int id, void** p) override {
    if (id < 0 || id >= max_inst_ || !slots_[id].live || !p) return fail("state_ptr: bad id");
    *p = h_inst_[id].A;
    return JX_OK;
  }

  // Tiles per resident wave that the per-round band heights aim for; 0 keeps every
  // instance at its cap.  Kernel arguments change, so the graph is captured again.
  int set_fill(int fill) override {
    if (fill < 0 || fill > 64) return fail("fill must be in [0, 64]");
    fill_ = fill;
    drop_graph();
    return JX_OK;
  }

  int set_graph(int on, int rounds) override {
    if (rounds < 1 || rounds > 64) return fail("graph rounds must be in [1, 64]");
    use_graph_ = on != 0;
    graph_rounds_ = rounds;
    drop_graph();
    return JX_OK;
  }

 private:
  struct Slot {
    bool live = false;
    size_t off = 0, bytes = 0;
    jx_run_params p{};
    long long since = 0;   // index of the first launch the instance takes part in
  };

  static size_t align(size_t b) { return (b + 255) & ~size_t(255); }
  // Kick sums of a forced instance: batch_kick_sums (7 in FP32, 3 in FP64) doubles per
  // (58-column strip, 8-row group).
  static size_t group_bytes(int nx, int ny) {
    return (size_t)batch_kick_sums<T>() * sizeof(double) * ((nx + kTileCols - 1) / kTileCols) *
           ((ny + kGroupRows - 1) / kGroupRows);
  }
  static constexpr size_t smem() { return (size_t)4 * kBatchSlice * 64 * sizeof(T); }

  int fail(const std::string& m, int code = JX_EINVAL) {
    err = m;
    return code;
  }

  BatchTabs<T> tabs() const {
    return BatchTabs<T>{d_inst_, d_ctrl_, d_meta_, d_cnt_, d_lists_, d_flags_,
                        max_inst_, nwaves_, fill_};
  }

  long long alloc(size_t need) {
    need = align(need);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= need) {
        const size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > need) free_[off + need] = sz - need;
        return (long long)off;
      }
    }
    return -1;
  }

  void release(size_t off, size_t sz) {
    sz = align(sz);
    auto it = free_.emplace(off, sz).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }

  // Uploads the instance table and the number of slots in use (the graph reads both from
  // device memory, so it stays valid across adds and removes).
  int upload_tables() {
    int top = 0;
    for (int k = 0; k < max_inst_; ++k)
      if (slots_[k].live) top = k + 1;
    if (hipMemcpyAsync(d_inst_, h_inst_.data(), sizeof(BatchInst<T>) * max_inst_,
                       hipMemcpyHostToDevice, stream_) != hipSuccess ||
        hipMemcpyAsync(d_meta_, &top, sizeof(int), hipMemcpyHostToDevice, stream_) !=
            hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess)
      return fail("table upload failed");
    return JX_OK;
  }

  // Which optional launches a round needs: the kick reduction (FP64: and the second-moment
  // pass and its reduction) while a forced instance may still step (kKickRound), the kick
  // pass while a forced instance without Smagorinsky may (kApplyRound; forced Smagorinsky
  // instances are kicked inside the timestep pass), the Smagorinsky timestep while an
  // instance with Smagorinsky may (kSmagRound).  An instance counts until a poll has seen it
  // done (done is final until the slot is reused).  Without such instances those launches'
  // work lists are empty, so leaving them out changes nothing but the round's fixed cost.
  static constexpr int kKickRound = 1, kSmagRound = 2, kApplyRound = 4, kRoundKinds = 8;
  int round_kind() const {
    int kind = 0;
    for (int k = 0; k < max_inst_; ++k)
      if (slots_[k].live && !done_seen_[k]) {
        if (h_inst_[k].forced) kind |= kKickRound;
        if (h_inst_[k].smag_on) kind |= kSmagRound;
        if (h_inst_[k].forced && !h_inst_[k].smag_on) kind |= kApplyRound;
      }
    return kind;
  }

  void launch_round(int kind) {
    const BatchTabs<T> tb = tabs();
    const dim3 g(grid_), blk(kBatchBlock);
    const dim3 gi((max_inst_ + 3) / 4);   // one wave per instance slot
    hipLaunchKernelGGL(batch_control_kernel<T>, dim3(1), dim3(kCtrlBlock), 0, stream_, tb);
    hipLaunchKernelGGL(batch_stage_kernel<T>, g, blk, smem(), stream_, tb);
    if (kind & kKickRound) {
      // FP32: mean force and scale from the final stage's one-pass sums; FP64: mean force,
      // then the second-moment pass and the scale.
      hipLaunchKernelGGL(batch_kick_reduce_kernel<T>, gi, dim3(256), 0, stream_, tb, 0);
      if (batch_kick_sums<T>() != 7) {
        hipLaunchKernelGGL(batch_kick_tiles_kernel<T>, g, blk, 0, stream_, tb, 0);
        hipLaunchKernelGGL(batch_kick_reduce_kernel<T>, gi, dim3(256), 0, stream_, tb, 1);
      }
    }
    if (kind & kApplyRound)
      hipLaunchKernelGGL(batch_kick_tiles_kernel<T>, g, blk, 0, stream_, tb, 1);
    if (kind & kSmagRound)
      hipLaunchKernelGGL(batch_timestep_kernel<T>, g, blk, 0, stream_, tb, -1);
  }

  // One graph per round kind, captured on first use.
  int capture(int kind) {
    if (hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("graph capture could not start");
    for (int k = 0; k < graph_rounds_; ++k) launch_round(kind);
    if (hipStreamEndCapture(stream_, &graph_[kind]) != hipSuccess || !graph_[kind])
      return fail("graph capture failed");
    if (hipGraphInstantiate(&graph_exec_[kind], graph_[kind], nullptr, nullptr, 0) !=
        hipSuccess)
      return fail("graph instantiation failed");
    return JX_OK;
  }

  void drop_graph() {
    for (int k = 0; k < kRoundKinds; ++k) {
      if (graph_exec_[k]) hipGraphExecDestroy(graph_exec_[k]);
      if (graph_[k]) hipGraphDestroy(graph_[k]);
      graph_exec_[k] = nullptr;
      graph_[k] = nullptr;
    }
  }

  int device_;
  int max_inst_;
  long long pool_bytes_;
  hipStream_t stream_ = nullptr;
  int cus_ = 0, grid_ = 1, nwaves_ = 4;
  int fill_ = 2;   // tiles per resident wave that the band heights aim for
  BatchInst<T>* d_inst_ = nullptr;
  BatchCtrl* d_ctrl_ = nullptr;
  int* d_meta_ = nullptr;
  unsigned* d_cnt_ = nullptr;
  int* d_lists_ = nullptr;
  int* d_flags_ = nullptr;
  void* d_pool_ = nullptr;
  // FP32 forcing render: per-slot mode-group tables (bases of up to kMaxGroupModes modes)
  // and the scratch they are computed in.
  static constexpr int kMaxGroupModes = 256;
  static constexpr int kGroupTable = 2 + 3 * kMaxGroupModes;
  int* d_groups_ = nullptr;
  int* d_gscratch_ = nullptr;
  BatchCtrl* h_ctrl_ = nullptr;
  int* h_meta_ = nullptr;
  // Pinned snapshots of the control table and counters taken at the end of each launch,
  // two deep, with their events (poll_launched reads the oldest unpolled one).
  BatchCtrl* h_snap_[2] = {};
  int* h_snap_meta_[2] = {};
  hipEvent_t snap_ev_[2] = {};
  int snap_head_ = 0, snap_tail_ = 0, snap_count_ = 0;
  long long snap_seq_[2] = {};   // the launch index each snapshot was taken at
  long long launches_ = 0;       // launches so far
  std::vector<BatchInst<T>> h_inst_;
  std::vector<Slot> slots_;
  std::vector<char> done_seen_;   // per slot: a poll has seen the instance done
  std::map<size_t, size_t> free_;
  bool use_graph_ = true;
  int graph_rounds_ = 6;
  hipGraph_t graph_[kRoundKinds] = {};
  hipGraphExec_t graph_exec_[kRoundKinds] = {};
};

}  // namespace jx

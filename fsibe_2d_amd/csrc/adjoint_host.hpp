// adjoint_host.hpp -- host side of the discrete adjoint (fsibe_adjoint_*),
// part of fsibe_hip.hip's translation unit (included at its end).
//
// Objective (no reference counterpart): J = (1/T) sum_n (-Force(1)^n) dt1^n
// over nsteps steps of a prescribed-motion simulation, T = sum dt1,
// Force(1) the cal_drag force
// (NavierStokes.f90:567-575); parameters: the 8 coefficients of motion 601
// (two harmonics of heave and of pitch).  Gradient by a reverse sweep over the
// adjoint kernels of kernels_adjoint.hpp.  The steps are cut into units
// (segments) of `chk` steps.  A unit is reversed by recomputing its steps
// from the state at its start while recording what the adjoint needs (the
// tape), then running its adjoint steps backwards.  The states at unit starts
// come from snapshot slots placed by a binomial checkpointing schedule
// (rev_plan): with a slot for every unit start, that is one forward sweep
// storing each unit's start; with fewer slots, the sweep advances again from
// the nearest stored state, and the number of recomputed units is the least
// any schedule with that many slots can take.  Memory: the slots plus one
// tape of chk steps.

static void adj_copy_d(hipStream_t st, const double *src, double *dst, size_t n) {
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256);
  ak_copy_d<<<blocks, 256, 0, st>>>(src, dst, (long long)n);
}
static void adj_copy_i(hipStream_t st, const int *src, int *dst, size_t n) {
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256);
  ak_copy_i<<<blocks, 256, 0, st>>>(src, dst, (long long)n);
}
static void adj_zero(hipStream_t st, double *dst, size_t n) {
  if (n == 0) return;
  const int blocks = (int)std::min<size_t>(1024, (n + 255) / 256);
  ak_zero<<<blocks, 256, 0, st>>>(dst, (long long)n);
}

struct AdjTape {
  int cap = 0, n = 0;
  size_t NH = 0, N = 0;
  double2 *A = nullptr, *B = nullptr;
  double *u1 = nullptr, *u2 = nullptr, *m = nullptr, *vx = nullptr, *vy = nullptr;
  int *win = nullptr;
  DevScalars *sc = nullptr;
  struct Rec {
    Box mbox;
    Window win;
    double tau = 0.0, dt1 = 0.0;
    int explicit_step = 0;
  };
  std::vector<Rec> rec;
};

// after create_mask: the mask box (values, velocities, winning primitives),
// the cal_drag window and the time of the kinematics the beam arrays hold
static void tape_mask(fsibe_sim *s) {
  AdjTape &t = *s->tape;
  if (t.n >= t.cap) return;
  const int j = t.n;
  AdjTape::Rec &r = t.rec[j];
  r.mbox = s->mbox;
  r.win = s->win[0];  // one beam: the adjoint runs iBeam = 1 only
  r.tau = s->beam_time;
  const size_t nb = s->mbox.empty() ? 0 : (size_t)s->mbox.w() * s->mbox.h();
  adj_copy_d(s->stream, s->mask_m, t.m + j * t.N, nb);
  adj_copy_d(s->stream, s->mask_vx, t.vx + j * t.N, nb);
  adj_copy_d(s->stream, s->mask_vy, t.vy + j * t.N, nb);
  if (s->mask_win) adj_copy_i(s->stream, s->mask_win, t.win + j * t.N, nb);
}

// after the nonlinear-term kernel: A, B (its column inputs), u of this step
// (it is overwritten by the velocity kernel) and the step's scalars
static int tape_fluid(fsibe_sim *s, bool explicit_step) {
  AdjTape &t = *s->tape;
  if (t.n >= t.cap) return set_err(FSIBE_ESTATE, "adjoint tape full");
  const int j = t.n;
  adj_copy_d(s->stream, reinterpret_cast<const double *>(s->T[0]), reinterpret_cast<double *>(t.A + j * t.NH),
             2 * t.NH);
  adj_copy_d(s->stream, reinterpret_cast<const double *>(s->T[1]), reinterpret_cast<double *>(t.B + j * t.NH),
             2 * t.NH);
  adj_copy_d(s->stream, s->u1, t.u1 + j * t.N, t.N);
  adj_copy_d(s->stream, s->u2, t.u2 + j * t.N, t.N);
  adj_copy_d(s->stream, reinterpret_cast<const double *>(s->dscal), reinterpret_cast<double *>(t.sc + j),
             sizeof(DevScalars) / sizeof(double));
  HIPCHK(hipGetLastError());
  t.rec[j].explicit_step = explicit_step ? 1 : 0;
  t.n = j + 1;
  return FSIBE_OK;
}

// the dynamic state of a simulation between two steps
struct SimSnap {
  char *dev = nullptr;
  double time = 0, dt0 = 0, dt1 = 0, beam_time = 0, jsum_f = 0, jsum_t = 0, drag_unst_old = 0, lift_unst_old = 0;
  int it = 0, n0 = 0, n1 = 1, slot = 0;
  bool startup_done = false, have_beam = false, unst_ready = false;
  std::vector<double> bx, by, bvx, bvy, btheta, btheta_dot;
};

static size_t snap_bytes(const fsibe_sim *s) {
  const size_t NH = (size_t)s->H * s->ny, N = (size_t)s->nx * s->ny;
  return sizeof(double2) * 4 * NH + sizeof(double) * (3 * N + s->ny) + sizeof(DevScalars);
}

// copy (to = true: sim -> snap) the device state, in the order W0 W1 NL0 NL1 u1 u2 vor_mean umax dscal
static void snap_dev(fsibe_sim *s, SimSnap &q, bool to) {
  const size_t NH = (size_t)s->H * s->ny, N = (size_t)s->nx * s->ny;
  double *p = reinterpret_cast<double *>(q.dev);
  double *fld[9] = {reinterpret_cast<double *>(s->W[0]), reinterpret_cast<double *>(s->W[1]),
                    reinterpret_cast<double *>(s->NL[0]), reinterpret_cast<double *>(s->NL[1]),
                    s->u1, s->u2, s->vor_mean, s->umax_part, reinterpret_cast<double *>(s->dscal)};
  const size_t len[9] = {2 * NH, 2 * NH, 2 * NH, 2 * NH, N, N, N, (size_t)s->ny, sizeof(DevScalars) / sizeof(double)};
  for (int k = 0; k < 9; ++k) {
    if (to) adj_copy_d(s->stream, fld[k], p, len[k]);
    else adj_copy_d(s->stream, p, fld[k], len[k]);
    p += len[k];
  }
}

static int snap_save(fsibe_sim *s, SimSnap &q) {
  snap_dev(s, q, true);
  HIPCHK(hipGetLastError());
  q.time = s->time; q.dt0 = s->dt0; q.dt1 = s->dt1; q.beam_time = s->beam_time;
  q.jsum_f = s->jsum_f; q.jsum_t = s->jsum_t; q.drag_unst_old = s->drag_unst_old[0];
  q.lift_unst_old = s->lift_unst_old[0];
  q.it = s->it; q.n0 = s->n0; q.n1 = s->n1; q.slot = s->slot;
  q.startup_done = s->startup_done; q.have_beam = s->have_beam; q.unst_ready = s->unst_ready[0];
  q.bx = s->bx; q.by = s->by; q.bvx = s->bvx; q.bvy = s->bvy; q.btheta = s->btheta; q.btheta_dot = s->btheta_dot;
  return FSIBE_OK;
}

static int snap_restore(fsibe_sim *s, SimSnap &q) {
  HIPCHK(hipStreamSynchronize(s->stream));
  snap_dev(s, q, false);
  HIPCHK(hipGetLastError());
  s->time = q.time; s->dt0 = q.dt0; s->dt1 = q.dt1; s->beam_time = q.beam_time;
  s->jsum_f = q.jsum_f; s->jsum_t = q.jsum_t; s->drag_unst_old[0] = q.drag_unst_old;
  s->lift_unst_old[0] = q.lift_unst_old;
  s->it = q.it; s->n0 = q.n0; s->n1 = q.n1; s->slot = q.slot;
  s->startup_done = q.startup_done; s->have_beam = q.have_beam; s->unst_ready[0] = q.unst_ready;
  s->bx = q.bx; s->by = q.by; s->bvx = q.bvx; s->bvy = q.bvy; s->btheta = q.btheta; s->btheta_dot = q.btheta_dot;
  s->res_pending = false;
  s->slot_fluid[0] = s->slot_fluid[1] = false;
  s->slot_forces[0] = s->slot_forces[1] = false;
  s->k1_ready = false;
  s->k2_ready = false;
  s->upload_pending = 0;
  return FSIBE_OK;
}

// ---- binomial checkpointing ------------------------------------------------
// Reversing l units with the state of unit 0 always at hand and s snapshot
// slots: the least number of units advanced (run forward without a tape,
// the first sweep included) is
//     t(l, s) = r l - B(s + 2, r - 1),   B(c, r) = (c + r)! / (c! r!),
// with r the least integer such that B(s + 1, r) >= l (Griewank, "Achieving
// logarithmic growth of temporal and spatial complexity in reverse
// automatic differentiation", Optim. Methods Softw. 1 (1992); Griewank and
// Walther, ACM TOMS 26 (2000), Algorithm 799 "revolve": s + 1 checkpoints
// there, counting the one of the initial state).  It satisfies the
// recursion t(l, 0) = l (l - 1) / 2, t(1, s) = 0,
//     t(l, s) = min over 1 <= k < l of  k + t(l - k, s - 1) + t(k, s)
// (advance k units, store the state there, reverse the l - k units above it
// with one slot fewer, then the k units below with the slot free again),
// and it is convex in l, so the objective of that minimum is convex in k
// and its least minimiser is found by bisection on its forward difference.

// t(l, s); the binomials B(s+1, r) grow term by term (B(c, r) = B(c, r-1)
// (c + r) / r, exact in 128 bits and saturated at 2^62)
static long long rev_cost(long long l, int s) {
  if (l <= 1) return 0;
  if (s <= 0) return l * (l - 1) / 2;
  const __int128 cap = (__int128)1 << 62;
  const int c = s + 1;
  long long r = 0;
  __int128 b = 1, b2 = 1;  // B(c, r) and B(c + 1, r - 1) (B(c + 1, -1) unused: l > 1 makes r >= 1)
  while (b < l) {
    ++r;
    b = b * (c + r) / r;
    if (r >= 2) b2 = b2 * (c + 1 + (r - 1)) / (r - 1);
    if (b > cap) b = cap;
    if (b2 > cap) b2 = cap;
  }
  return r * l - (long long)b2;
}

// the least k in [1, l - 1] minimising k + t(l - k, s - 1) + t(k, s)
static long long rev_split(long long l, int s) {
  auto f = [&](long long k) { return k + rev_cost(l - k, s - 1) + rev_cost(k, s); };
  long long lo = 1, hi = l - 1;  // the answer is the first k with f(k+1) - f(k) >= 0, or l - 1
  while (lo < hi) {
    const long long mid = lo + (hi - lo) / 2;
    if (f(mid + 1) - f(mid) >= 0) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// one action of the schedule: restore the state of unit `unit` from slot
// `arg` (-1: the initial state), advance `arg` units from unit `unit`, store
// the state of unit `unit` in slot `arg`, or record and reverse unit `unit`
enum RevOp { REV_RESTORE = 0, REV_ADVANCE = 1, REV_SAVE = 2, REV_REVERSE = 3 };
struct RevAct {
  int op, unit, arg;
};

// reverse units [a, b), the state of unit a being in slot sa, with s free
// slots (ids taken from `free_ids`)
static void rev_plan_rec(long long a, long long b, int s, int sa, std::vector<int> &free_ids,
                         std::vector<RevAct> &out) {
  if (b - a > 1 && s >= b - a - 1) {
    // a slot for every unit start above a: one sweep storing each (the
    // schedule's splits would all be k = 1), then every unit in turn
    std::vector<int> ids((size_t)(b - a));
    ids[0] = sa;
    out.push_back({REV_RESTORE, (int)a, sa});
    for (long long u = a + 1; u < b; ++u) {
      ids[(size_t)(u - a)] = free_ids.back();
      free_ids.pop_back();
      out.push_back({REV_ADVANCE, (int)(u - 1), 1});
      out.push_back({REV_SAVE, (int)u, ids[(size_t)(u - a)]});
    }
    for (long long u = b - 1; u >= a; --u) {
      out.push_back({REV_RESTORE, (int)u, ids[(size_t)(u - a)]});
      out.push_back({REV_REVERSE, (int)u, 0});
      if (u > a) free_ids.push_back(ids[(size_t)(u - a)]);
    }
    return;
  }
  while (b - a > 1 && s > 0) {
    const long long k = rev_split(b - a, s);
    const int q = free_ids.back();
    free_ids.pop_back();
    out.push_back({REV_RESTORE, (int)a, sa});
    out.push_back({REV_ADVANCE, (int)a, (int)k});
    out.push_back({REV_SAVE, (int)(a + k), q});
    rev_plan_rec(a + k, b, s - 1, q, free_ids, out);
    free_ids.push_back(q);
    b = a + k;
  }
  for (long long i = b - 1; i >= a; --i) {  // no slot left: advance from a for each unit
    out.push_back({REV_RESTORE, (int)a, sa});
    if (i > a) out.push_back({REV_ADVANCE, (int)a, (int)(i - a)});
    out.push_back({REV_REVERSE, (int)i, 0});
  }
}

static std::vector<RevAct> rev_plan(int nunits, int nslots) {
  std::vector<RevAct> out;
  std::vector<int> free_ids;
  for (int q = nslots - 1; q >= 0; --q) free_ids.push_back(q);
  if (nunits > 0) rev_plan_rec(0, nunits, nslots, -1, free_ids, out);
  return out;
}

// The schedule as integer triples (op, unit, arg) for a host to inspect:
// ops[3*i .. 3*i+2] of action i, *n_acts actions (FSIBE_ESIZE with the count
// when cap is short), *advances the number of units it advances.
extern "C" int fsibe_adjoint_plan(int nunits, int nslots, int *ops, int cap, int *n_acts, long long *advances) {
  if (nunits < 1 || nslots < 0 || nunits > (1 << 24)) return set_err(FSIBE_EARG, "adjoint plan: bad sizes");
  nslots = std::min(nslots, nunits - 1);
  const std::vector<RevAct> pl = rev_plan(nunits, nslots);
  long long adv = 0;
  for (const RevAct &x : pl)
    if (x.op == REV_ADVANCE) adv += x.arg;
  if (n_acts) *n_acts = (int)pl.size();
  if (advances) *advances = adv;
  if (!ops) return FSIBE_OK;
  if ((size_t)cap < pl.size()) return set_err(FSIBE_ESIZE, "adjoint plan: %zu actions", pl.size());
  for (size_t i = 0; i < pl.size(); ++i) {
    ops[3 * i] = pl[i].op;
    ops[3 * i + 1] = pl[i].unit;
    ops[3 * i + 2] = pl[i].arg;
  }
  return FSIBE_OK;
}

// the least number of units any schedule advances (t(l, s) above)
extern "C" long long fsibe_adjoint_plan_cost(int nunits, int nslots) {
  if (nunits < 1 || nslots < 0) return -1;
  return rev_cost(nunits, std::min(nslots, nunits - 1));
}

struct fsibe_adjoint {
  fsibe_sim *s = nullptr;
  int nsteps = 0, chk = 0, nseg = 0, nslots = 0;
  long long steps_advanced = 0, steps_recorded = 0;
  SimSnap init;
  std::vector<SimSnap> snaps;
  AdjTape tape;
  double2 *lW = nullptr, *lNL = nullptr, *T[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double *lu1 = nullptr, *lu2 = nullptr, *lm = nullptr, *lvx = nullptr, *lvy = nullptr;
  double *part = nullptr, *h_part = nullptr;
  int part_blocks = 0;
  double *d_nodes = nullptr;
  std::vector<char *> allocs;
  double ms_forward = 0.0, ms_gradient = 0.0, ms_reverse = 0.0;
};

static char *adj_alloc(fsibe_adjoint *a, size_t bytes) {
  char *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) return nullptr;
  a->allocs.push_back(p);
  return p;
}

extern "C" int fsibe_adjoint_destroy(fsibe_adjoint *a) {
  if (!a) return FSIBE_OK;
  if (a->s) (void)hipStreamSynchronize(a->s->stream);
  for (char *p : a->allocs) (void)hipFree(p);
  if (a->h_part) (void)hipHostFree(a->h_part);
  if (a->s) fsibe_sim_destroy(a->s);
  delete a;
  return FSIBE_OK;
}

extern "C" int fsibe_adjoint_create_binomial(const fsibe_params *p, const double *vortk, int nsteps, int chk,
                                             int nsnap, fsibe_adjoint **out) {
  if (!p || !out || nsteps < 1 || chk < 1) return set_err(FSIBE_EARG, "adjoint: bad argument");
  *out = nullptr;
  if (p->iFLUSI != 0 || p->iBeam != 1 || p->iMotion != 601)
    return set_err(FSIBE_EUNSUPPORTED, "adjoint: needs prescribed motion 601 of one beam (iFLUSI=0, iBeam=1)");
  if (p->iSharpTrailing != 0 || p->iSponge == 10 || p->iSponge == 4 || p->iSponge == 22 || p->iMeanVelocity == 11)
    return set_err(FSIBE_EUNSUPPORTED,
                   "adjoint: sharp trailing edge, iSponge 4/10/22 and iMeanVelocity 11 are not differentiated");
  if (!(p->dt_fixed > 0.0))
    return set_err(FSIBE_EUNSUPPORTED, "adjoint: needs a fixed time step (dt_fixed > 0); dt1 is not differentiated");
  fsibe_adjoint *a = new fsibe_adjoint();
  auto fail = [&](int rc) { fsibe_adjoint_destroy(a); return rc; };
  int rc = fsibe_sim_create(p, &a->s);
  if (rc != FSIBE_OK) return fail(rc);
  fsibe_sim *s = a->s;
  if (vortk) {
    rc = fsibe_set_vortk(s, vortk);
    if (rc != FSIBE_OK) return fail(rc);
  }
  a->nsteps = nsteps;
  a->chk = std::min(chk, nsteps);
  a->nseg = (nsteps + a->chk - 1) / a->chk;
  a->nslots = nsnap < 0 ? a->nseg - 1 : std::min(nsnap, a->nseg - 1);
  const size_t NH = (size_t)s->H * s->ny, N = (size_t)s->nx * s->ny;
  const int cap = a->chk;
  AdjTape &t = a->tape;
  t.cap = cap; t.NH = NH; t.N = N;
  t.rec.resize(cap);
  bool ok = true;
  auto D = [&](size_t bytes) { char *q = adj_alloc(a, bytes); ok = ok && q; return q; };
  if (hipMalloc(&s->mask_win, sizeof(int) * N) != hipSuccess) ok = false;  // freed with the simulation
  t.A = reinterpret_cast<double2 *>(D(sizeof(double2) * NH * cap));
  t.B = reinterpret_cast<double2 *>(D(sizeof(double2) * NH * cap));
  t.u1 = reinterpret_cast<double *>(D(sizeof(double) * N * cap));
  t.u2 = reinterpret_cast<double *>(D(sizeof(double) * N * cap));
  t.m = reinterpret_cast<double *>(D(sizeof(double) * N * cap));
  t.vx = reinterpret_cast<double *>(D(sizeof(double) * N * cap));
  t.vy = reinterpret_cast<double *>(D(sizeof(double) * N * cap));
  t.win = reinterpret_cast<int *>(D(sizeof(int) * N * cap));
  t.sc = reinterpret_cast<DevScalars *>(D(sizeof(DevScalars) * cap));
  a->lW = reinterpret_cast<double2 *>(D(sizeof(double2) * NH));
  a->lNL = reinterpret_cast<double2 *>(D(sizeof(double2) * NH));
  for (int k = 0; k < 6; ++k) a->T[k] = reinterpret_cast<double2 *>(D(sizeof(double2) * NH));
  a->lu1 = reinterpret_cast<double *>(D(sizeof(double) * N));
  a->lu2 = reinterpret_cast<double *>(D(sizeof(double) * N));
  a->lm = reinterpret_cast<double *>(D(sizeof(double) * N));
  a->lvx = reinterpret_cast<double *>(D(sizeof(double) * N));
  a->lvy = reinterpret_cast<double *>(D(sizeof(double) * N));
  a->part_blocks = 256;
  a->part = reinterpret_cast<double *>(D(sizeof(double) * 4 * a->part_blocks * cap));
  a->d_nodes = reinterpret_cast<double *>(D(sizeof(double) * std::max(s->ns, 1)));
  a->init.dev = D(snap_bytes(s));
  a->snaps.resize(a->nslots);
  for (auto &q : a->snaps) q.dev = D(snap_bytes(s));
  if (!ok) return fail(set_err(FSIBE_EHIP, "adjoint: device allocation failed"));
  if (hipHostMalloc((void **)&a->h_part, sizeof(double) * 4 * a->part_blocks * cap, hipHostMallocDefault) != hipSuccess)
    return fail(set_err(FSIBE_EHIP, "adjoint: host allocation failed"));
  // node arc lengths: integrate_position's prefix sums for theta = 0
  std::vector<double> sn(std::max(s->ns, 1), 0.0);
  for (int i = 1; i < s->ns; ++i) sn[i] = sn[i - 1] + 1.0 * s->ds;
  if (hipMemcpy(a->d_nodes, sn.data(), sizeof(double) * sn.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail(set_err(FSIBE_EHIP, "adjoint: upload failed"));
  rc = snap_save(s, a->init);
  if (rc == FSIBE_OK && hipStreamSynchronize(s->stream) != hipSuccess) rc = set_err(FSIBE_EHIP, "adjoint: sync failed");
  if (rc != FSIBE_OK) return fail(rc);
  *out = a;
  return FSIBE_OK;
}

extern "C" int fsibe_adjoint_create(const fsibe_params *p, const double *vortk, int nsteps, int chk,
                                    fsibe_adjoint **out) {
  return fsibe_adjoint_create_binomial(p, vortk, nsteps, chk, -1, out);
}

// the initial state with the motion coefficients c
static int adj_begin(fsibe_adjoint *a, const double coef[8]) {
  fsibe_sim *s = a->s;
  for (int k = 0; k < 8; ++k) s->p.motion_coef[k] = coef[k];
  CHK(snap_restore(s, a->init));
  s->have_beam = false;
  init_beam_host(s);
  s->jsum_f = 0.0;
  s->jsum_t = 0.0;
  return FSIBE_OK;
}

// first step of unit k
static int unit_start(const fsibe_adjoint *a, int k) { return std::min(k * a->chk, a->nsteps); }

// run the steps of units [k, k + n) forward from the state of unit k
static int adj_advance(fsibe_adjoint *a, int k, int n) {
  fsibe_sim *s = a->s;
  const int first = unit_start(a, k), end = unit_start(a, k + n);
  fsibe_step_out o;
  CHK(fsibe_run(s, end - first, &o));
  if (s->it + 1 < end) return set_err(FSIBE_ESTATE, "adjoint: the run reached Time_end");
  a->steps_advanced += end - first;
  return FSIBE_OK;
}

static int adj_forward(fsibe_adjoint *a, double *J) {
  fsibe_sim *s = a->s;
  CHK(adj_advance(a, 0, a->nseg));
  if (!(s->jsum_t > 0.0)) return set_err(FSIBE_ESTATE, "adjoint: no step taken");
  *J = s->jsum_f / s->jsum_t;
  return FSIBE_OK;
}

extern "C" int fsibe_adjoint_objective(fsibe_adjoint *a, const double coef[8], double *J) {
  if (!a || !coef || !J) return set_err(FSIBE_EARG, "null argument");
  const double t0 = now_ms();
  a->steps_advanced = 0;
  a->steps_recorded = 0;
  CHK(adj_begin(a, coef));
  CHK(adj_forward(a, J));
  a->ms_forward = now_ms() - t0;
  return FSIBE_OK;
}

// d(LEy, alpha, hdot, alpha_t)/d coef at time t for motion 601 (kinematics.hpp)
static void motion601_jac(const fsibe_params &p, double t, double jac[4][8]) {
  const double pi = 4.0 * atan(1.0), w0 = 2.0 * pi * p.heave_freq;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 8; ++j) jac[i][j] = 0.0;
  for (int m = 1; m <= 2; ++m) {
    const double wm = w0 * m, sn = sin(wm * t), cs = cos(wm * t);
    const int ja = 2 * (m - 1), jb = ja + 1;
    jac[0][ja] = sn;             // LEy = y0 + sum a sin + b (cos - 1)
    jac[0][jb] = cs - 1.0;
    jac[2][ja] = wm * cs;        // hdot
    jac[2][jb] = -(wm * sn);
    jac[1][4 + ja] = sn;         // alpha = AngleBeam pi/180 + sum ...
    jac[1][4 + jb] = cs - 1.0;
    jac[3][4 + ja] = wm * cs;    // alpha_t
    jac[3][4 + jb] = -(wm * sn);
  }
}

static int adj_reverse_step(fsibe_adjoint *a, int j, const double *u1n, double cJ) {
  fsibe_sim *s = a->s;
  AdjTape &t = a->tape;
  const AdjTape::Rec &r = t.rec[j];
  MaskBox mb;
  mb.x0 = r.mbox.x0; mb.x1 = r.mbox.x1; mb.y0 = r.mbox.y0; mb.y1 = r.mbox.y1; mb.w = r.mbox.w();
  if (r.mbox.empty()) { mb.x0 = 0; mb.x1 = -1; mb.y0 = 0; mb.y1 = -1; mb.w = 0; }
  mb.m = t.m + j * t.N; mb.vx = t.vx + j * t.N; mb.vy = t.vy + j * t.N;
  double2 *Ta = a->T[0], *Tb = a->T[1], *Tc = a->T[2], *Td = a->T[3], *Te = a->T[4], *Tf = a->T[5];
  const size_t nbox = r.mbox.empty() ? 0 : (size_t)r.mbox.w() * r.mbox.h();
  adj_zero(s->stream, a->lm, nbox);
  adj_zero(s->stream, a->lvx, nbox);
  adj_zero(s->stream, a->lvy, nbox);
  CHK(with_rows<2>(s->px, [&](auto F, int nt) -> int {
    KLAUNCH(ak_rows_vel<decltype(F)>, s->ny, nt, sizeof(double2) * 2 * padlen(s->H), s->stream, a->lu1, a->lu2,
            u1n, mb, r.win, cJ, a->lm, a->lvx, a->lvy, Ta, Tb, F, s->twx, s->sc);
    return FSIBE_OK;
  }));
  CHK(with_cols<2>(s->py, s->H, pick_tx(s->H, s->ny, 2, 1), [&](auto F, int nt, int TX) -> int {
    KLAUNCH(ak_cols_ab2<decltype(F)>, s->H / TX, nt, sizeof(double2) * 2 * TX * padlen(s->ny), s->stream, Ta, Tb,
            a->lW, a->lNL, Tc, Td, t.sc + j, r.explicit_step, F, s->twy, s->sc);
    return FSIBE_OK;
  }));
  CHK(with_rows<6>(s->px, [&](auto F, int nt) -> int {
    KLAUNCH(ak_rows_nl<decltype(F)>, s->ny, nt, sizeof(double2) * 6 * padlen(s->H), s->stream, Tc, Td,
            t.A + j * t.NH, t.B + j * t.NH, t.u1 + j * t.N, t.u2 + j * t.N,
            s->sponge_on ? s->ms : nullptr, mb, (int)s->penal_on, a->lu1, a->lu2, a->lm, a->lvx, a->lvy, Te, Tf,
            F, s->twx, s->sc);
    return FSIBE_OK;
  }));
  CHK(with_cols<2>(s->py, s->H, pick_tx(s->H, s->ny, 2, 1), [&](auto F, int nt, int TX) -> int {
    KLAUNCH(ak_cols_w0<decltype(F)>, s->H / TX, nt, sizeof(double2) * 2 * TX * padlen(s->ny), s->stream, Te, Tf,
            a->lW, F, s->twy, s->sc);
    return FSIBE_OK;
  }));
  double *part = a->part + (size_t)j * 4 * a->part_blocks;
  if (nbox > 0) {
    double alpha, at, att, LE[6];
    mouvement_host(s->p, s->p.iMotion, s->p.x0, s->p.y0, s->p.AngleBeam, 0.0, r.tau, alpha, at, att, LE);
    MaskConsts mc;
    mc.dx = s->dx; mc.dy = s->dy; mc.t = s->p.t_beam; mc.h = s->p.N_smooth * std::max(s->dx, s->dy);
    mc.gerf = s->gerf; mc.inv_eps_unused = 0.0; mc.ds = s->ds; mc.pi = s->pi; mc.sharp = s->sharp;
    mc.eps = s->p.eps;
    RigidBeam rb;
    rb.LEx = LE[0]; rb.LEy = LE[1]; rb.alpha = alpha; rb.hdot = LE[3]; rb.at = at;
    rb.s = a->d_nodes; rb.ns = s->ns; rb.off = s->p.iCylinder == 1 ? 1 : 0;
    ak_mask<<<a->part_blocks, 256, 0, s->stream>>>(t.win + j * t.N, a->lm, a->lvx, a->lvy, mc, rb, r.mbox.x0,
                                                  r.mbox.y0, r.mbox.w(), r.mbox.h(), part);
    HIPCHK(hipGetLastError());
  } else {
    adj_zero(s->stream, part, 4 * (size_t)a->part_blocks);
  }
  return FSIBE_OK;
}

// record unit k (its steps recomputed from the state of its start, which the
// simulation holds) and run its adjoint steps backwards; the motion
// parameters' partial sums go into g.  The first unit recorded is the last
// one: its run ends the forward sweep and fixes J and T.
static int adj_record_reverse(fsibe_adjoint *a, int k, double *J, double *T, double g[8]) {
  fsibe_sim *s = a->s;
  const int first = unit_start(a, k), len = unit_start(a, k + 1) - first;
  AdjTape &t = a->tape;
  t.n = 0;
  s->tape = &t;
  int rc = FSIBE_OK;
  // pipelined as fsibe_run's prescribed-motion loop: the host prepares
  // step j+1 while step j runs and reads each step's results one step late
  for (int j = 0; j < len && rc == FSIBE_OK; ++j) {
    rc = run_one(s, true, false);
    if (rc == FSIBE_OK) {
      t.rec[j].dt1 = s->dt1;
      if (t.n != j + 1) rc = set_err(FSIBE_ESTATE, "adjoint: step %d was not recorded", first + j);
      else if (s->dt1 != s->p.dt_fixed)
        rc = set_err(FSIBE_EDOMAIN, "adjoint: step %d took dt1 = %g from the CFL bound or Time_end, not dt_fixed",
                     first + j, s->dt1);
    }
  }
  s->tape = nullptr;
  s->pipe = false;
  if (s->res_pending) {  // the last step's results (the reverse sweep needs none of them)
    if (hipStreamSynchronize(s->stream) != hipSuccess && rc == FSIBE_OK) rc = set_err(FSIBE_EHIP, "adjoint: sync failed");
    parse_results(s, s->slot ^ 1);
    s->res_pending = false;
  }
  CHK(rc);
  a->steps_recorded += len;
  if (!(*T > 0.0)) {
    if (k != a->nseg - 1) return set_err(FSIBE_ESTATE, "adjoint: the schedule reversed unit %d first", k);
    if (s->it + 1 < a->nsteps) return set_err(FSIBE_ESTATE, "adjoint: the run reached Time_end");
    if (!(s->jsum_t > 0.0)) return set_err(FSIBE_ESTATE, "adjoint: no step taken");
    *J = s->jsum_f / s->jsum_t;
    *T = s->jsum_t;
  }
  const double dxdy = s->dx * s->dy;
  for (int j = len - 1; j >= 0; --j) {
    const double *u1n = j == len - 1 ? s->u1 : t.u1 + (size_t)(j + 1) * t.N;
    CHK(adj_reverse_step(a, j, u1n, -(t.rec[j].dt1 * dxdy) / *T));
  }
  HIPCHK(hipMemcpyAsync(a->h_part, a->part, sizeof(double) * 4 * a->part_blocks * len, hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int j = len - 1; j >= 0; --j) {
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    const double *pp = a->h_part + (size_t)j * 4 * a->part_blocks;
    for (int b = 0; b < a->part_blocks; ++b)
      for (int i = 0; i < 4; ++i) q[i] += pp[4 * b + i];
    double jac[4][8];
    motion601_jac(s->p, t.rec[j].tau, jac);
    for (int c = 0; c < 8; ++c)
      for (int i = 0; i < 4; ++i) g[c] += q[i] * jac[i][c];
  }
  return FSIBE_OK;
}

extern "C" int fsibe_adjoint_gradient(fsibe_adjoint *a, const double coef[8], double *J, double grad[8]) {
  if (!a || !coef || !J || !grad) return set_err(FSIBE_EARG, "null argument");
  fsibe_sim *s = a->s;
  const double t0 = now_ms();
  double t1 = -1.0;
  a->steps_advanced = 0;
  a->steps_recorded = 0;
  CHK(adj_begin(a, coef));
  int cur = 0;  // the unit whose start the simulation's state is
  const size_t NH = (size_t)s->H * s->ny, N = (size_t)s->nx * s->ny;
  adj_zero(s->stream, reinterpret_cast<double *>(a->lW), 2 * NH);
  adj_zero(s->stream, reinterpret_cast<double *>(a->lNL), 2 * NH);
  adj_zero(s->stream, a->lu1, N);
  adj_zero(s->stream, a->lu2, N);
  double g[8] = {0, 0, 0, 0, 0, 0, 0, 0}, T = 0.0;
  for (const RevAct &x : rev_plan(a->nseg, a->nslots)) {
    switch (x.op) {
      case REV_RESTORE:
        if (cur != x.unit) {
          if (x.arg < 0) CHK(adj_begin(a, coef));
          else CHK(snap_restore(s, a->snaps[x.arg]));
          cur = x.unit;
        }
        break;
      case REV_ADVANCE:
        if (cur != x.unit) return set_err(FSIBE_ESTATE, "adjoint: schedule out of order");
        CHK(adj_advance(a, x.unit, x.arg));
        cur = x.unit + x.arg;
        break;
      case REV_SAVE:
        if (cur != x.unit) return set_err(FSIBE_ESTATE, "adjoint: schedule out of order");
        CHK(snap_save(s, a->snaps[x.arg]));
        break;
      case REV_REVERSE:
        if (cur != x.unit) return set_err(FSIBE_ESTATE, "adjoint: schedule out of order");
        if (t1 < 0.0) t1 = now_ms();
        CHK(adj_record_reverse(a, x.unit, J, &T, g));
        cur = x.unit + 1;
        break;
      default: break;
    }
  }
  for (int c = 0; c < 8; ++c) grad[c] = g[c];
  const double t2 = now_ms();
  a->ms_gradient = t2 - t0;
  a->ms_reverse = t2 - t1;
  return FSIBE_OK;
}

// the checkpoint slots and, for the last objective or gradient call, the
// steps advanced without a tape (a gradient's forward sweep included) and
// the steps recorded
extern "C" int fsibe_adjoint_stats(fsibe_adjoint *a, int *slots, long long *steps_advanced,
                                   long long *steps_recorded) {
  if (!a) return set_err(FSIBE_EARG, "null adjoint");
  if (slots) *slots = a->nslots;
  if (steps_advanced) *steps_advanced = a->steps_advanced;
  if (steps_recorded) *steps_recorded = a->steps_recorded;
  return FSIBE_OK;
}

extern "C" int fsibe_adjoint_timings(fsibe_adjoint *a, double *forward_ms, double *gradient_ms, double *reverse_ms) {
  if (!a) return set_err(FSIBE_EARG, "null adjoint");
  if (forward_ms) *forward_ms = a->ms_forward;
  if (gradient_ms) *gradient_ms = a->ms_gradient;
  if (reverse_ms) *reverse_ms = a->ms_reverse;
  return FSIBE_OK;
}

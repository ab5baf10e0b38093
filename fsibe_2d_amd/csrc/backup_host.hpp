// backup_host.hpp -- the reference's restart files, written from and read
// into a device-resident simulation; part of fsibe_hip.hip's translation
// unit (included at its end).
//
//   fsibe_backup_write  MakeRuntimeBackup (save_fields.f90:109-152)
//   fsibe_backup_read   init_fields, inicond = 2 (init_fields.f90:44-97)
//   fsibe_inicond_write the final vorticity file (time_step.f90:350-355)
//   fsibe_inicond_read  init_fields, inicond = 111 (init_fields.f90:109-128)
//
// Format: Fortran unformatted sequential records as the reference's build
// (ifort -r8) writes them: each record is a 4-byte little-endian byte count,
// the items in order, and the same count again; default REAL and real
// literals are 8 bytes, INTEGER and LOGICAL 4 (.true. = -1, read back as
// true when odd or nonzero).  Arrays are column-major (x fastest), which is
// the library's row-major [y][x] order; spectral arrays in the packed
// layout of cof_fftw33.f90.

// per-row max of u1^2 + u2^2, as the velocity kernel leaves it for the next
// step's CFL limit
__global__ void k_row_umax(const double *__restrict__ u1, const double *__restrict__ u2, double *__restrict__ part,
                           int nx, int ny) {
  __shared__ double red[4];
  const int y = blockIdx.x;
  double m = 0.0;
  for (int x = threadIdx.x; x < nx; x += blockDim.x) {
    const double a = u1[(size_t)y * nx + x], b = u2[(size_t)y * nx + x];
    m = fmax(m, a * a + b * b);
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[y] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

namespace {

struct FItem {
  const void *p;
  size_t n;
};

// one Fortran record; false on an I/O error or a record of 2 GiB or more
bool frec_write(FILE *f, std::initializer_list<FItem> items) {
  size_t len = 0;
  for (const FItem &it : items) len += it.n;
  if (len > 0x7fffffffu) return false;
  const int32_t m = (int32_t)len;
  if (fwrite(&m, 4, 1, f) != 1) return false;
  for (const FItem &it : items)
    if (it.n && fwrite(it.p, 1, it.n, f) != it.n) return false;
  return fwrite(&m, 4, 1, f) == 1;
}

// the next record into buf; false at end of file or on a malformed record
bool frec_read(FILE *f, std::vector<char> &buf) {
  int32_t m = 0, m2 = 0;
  if (fread(&m, 4, 1, f) != 1 || m < 0) return false;
  buf.resize((size_t)m);
  if (m && fread(buf.data(), 1, (size_t)m, f) != (size_t)m) return false;
  return fread(&m2, 4, 1, f) == 1 && m2 == m;
}

// sequential reader of one record's items
struct FCursor {
  const std::vector<char> &b;
  size_t o = 0;
  bool ok = true;
  explicit FCursor(const std::vector<char> &buf) : b(buf) {}
  void get(void *dst, size_t n) {
    if (!ok || o + n > b.size()) { ok = false; return; }
    memcpy(dst, b.data() + o, n);
    o += n;
  }
  bool done() const { return ok && o == b.size(); }
};

const int32_t kFTrue = -1;  // ifort's .true.

}  // namespace

// exp(-nu k^2 dt) in the packed layout, cal_vis.f90:18-34
static void vis_packed_host(const fsibe_sim *s, double dt, double *w) {
  const int nx = s->nx, ny = s->ny;
  const double coefx = -dt * s->p.nu * s->fp.scalex, coefy = -dt * s->p.nu * s->fp.scaley;
  for (int ky = 0; ky <= ny / 2 - 1; ++ky)
    for (int kx = 0; kx <= nx / 2 - 1; ++kx) {
      const double v = exp(coefx * (double)(kx * kx) + coefy * (double)(ky * ky));
      w[(size_t)(2 * ky) * nx + 2 * kx] = v;
      w[(size_t)(2 * ky) * nx + 2 * kx + 1] = v;
      w[(size_t)(2 * ky + 1) * nx + 2 * kx] = v;
      w[(size_t)(2 * ky + 1) * nx + 2 * kx + 1] = v;
    }
}

// a half-spectrum field in the packed layout, on the host
static int download_packed(fsibe_sim *s, const double2 *src, double *host) {
  const size_t N = (size_t)s->nx * s->ny;
  k_half_to_packed<<<grid1d((size_t)s->H * (s->ny / 2)), 256, 0, s->stream>>>(src, s->fb_p, s->nx, s->ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host, s->fb_p, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return FSIBE_OK;
}

static int upload_packed(fsibe_sim *s, const double *host, double2 *dst) {
  const size_t N = (size_t)s->nx * s->ny;
  HIPCHK(hipMemcpyAsync(s->fb_p, host, sizeof(double) * N, hipMemcpyHostToDevice, s->stream));
  k_packed_to_half<<<grid1d((size_t)s->H * (s->ny / 2)), 256, 0, s->stream>>>(s->fb_p, dst, s->nx, s->ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  return FSIBE_OK;
}

extern "C" int fsibe_backup_write(fsibe_sim *s, const char *path, int ivideo, double colorscale,
                                  int colorscale_done) {
  if (!s || !path) return set_err(FSIBE_EARG, "null argument");
  HIPCHK(hipStreamSynchronize(s->stream));
  const size_t N = (size_t)s->nx * s->ny;
  const int n0 = s->n0, n1 = s->n1;
  std::vector<double> vk(2 * N), nl(2 * N), wv(2 * N, 0.0), mk(3 * N), u(2 * N), vm(N);
  const bool restored = s->restored.size() == 9 * N;
  if (restored) {
    // no step since fsibe_backup_read: the file's own arrays
    const double *r = s->restored.data();
    std::copy(r, r + 2 * N, vk.begin());
    std::copy(r + 2 * N, r + 4 * N, nl.begin());
    std::copy(r + 4 * N, r + 6 * N, wv.begin());
    std::copy(r + 6 * N, r + 9 * N, mk.begin());
  }
  // vortk and nlk at the Fortran level indices 0, 1 the library also uses
  if (!restored) CHK(download_packed(s, s->W[n0], vk.data() + n0 * N));
  if (!restored) {
    CHK(download_packed(s, s->W[n1], vk.data() + n1 * N));
    CHK(download_packed(s, s->NL[n1], nl.data() + n1 * N));
    CHK(download_packed(s, s->NL[n0], nl.data() + n0 * N));
  }
  // workvis(n0) = vis(dt1) of the last step, workvis(n1) = vis(dt0) of the
  // step before (never set after the startup step alone); the reference's
  // nlk(n0) holds the previous nonlinear term times workvis of its step
  // (NavierStokes.f90:478-481), which the library applies on the fly
  if (!restored) {
    vis_packed_host(s, s->dt1, wv.data() + n0 * N);
    if (s->it >= 1) {
      vis_packed_host(s, s->dt0, wv.data() + n1 * N);
      double *a = nl.data() + n0 * N;
      const double *w = wv.data() + n1 * N;
      for (size_t i = 0; i < N; ++i) a[i] = a[i] * w[i];
    }
    CHK(box_to_full(s, s->mask_m, s->mbox, mk.data()));
    CHK(box_to_full(s, s->mask_vx, s->mbox, mk.data() + N));
    CHK(box_to_full(s, s->mask_vy, s->mbox, mk.data() + 2 * N));
  }
  HIPCHK(hipMemcpyAsync(u.data(), s->u1, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(u.data() + N, s->u2, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(vm.data(), s->vor_mean, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));

  FILE *f = fopen(path, "wb");
  if (!f) return set_err(FSIBE_EARG, "backup: cannot open %s for writing", path);
  const double chk = 123.0, t = s->time, dt1 = s->dt1;
  const int32_t in1 = n1, iv = ivideo, ns = s->ns, cd = colorscale_done ? kFTrue : 0;
  bool ok = frec_write(f, {{&t, 8}}) && frec_write(f, {{&chk, 8}}) &&
            frec_write(f, {{&in1, 4}, {&dt1, 8}, {vk.data(), 16 * N}, {nl.data(), 16 * N}, {wv.data(), 16 * N},
                           {mk.data(), 8 * N}, {mk.data() + N, 8 * N}, {mk.data() + 2 * N, 8 * N}, {&iv, 4},
                           {u.data(), 16 * N}}) &&
            frec_write(f, {{&ns, 4}}) && frec_write(f, {{&colorscale, 8}, {&cd, 4}});
  // six records per beam (save_fields.f90:132-139, do i = 1, iBeam)
  for (int ib = 0; ok && ib < s->nb; ++ib) {
    const size_t B = 8 * (size_t)ns, o = (size_t)ib * ns;
    const bool solid = s->p.iFLUSI == 1 && s->solid[ib];
    std::vector<double> zero(6 * (size_t)ns, 0.0);
    const ibes::Beam *b = solid ? &s->solid[ib]->b : nullptr;
    const double *pn = solid ? b->pressure_new.data() : s->pbeam.data() + o;
    const double *po = solid ? b->pressure_old.data() : s->pbeam.data() + o;
    const double *tn = solid ? b->tau_new.data() : s->tbeam.data() + o;
    const double *to = solid ? b->tau_old.data() : s->tbeam.data() + o;
    const double Ek = solid ? b->E_kinetic : 0.0, Ep = solid ? b->E_pot : 0.0, Ee = solid ? b->E_elastic : 0.0;
    double x0 = 0.0, y0 = 0.0, ph = 0.0;
    beam_placement(s->p, ib, x0, y0, ph);
    const double ab = s->p.AngleBeam, dto = solid ? b->dt_old : 0.0;
    const double dun = s->drag_unst_old[ib], lun = s->lift_unst_old[ib];
    const int32_t imv = s->p.iMotion, ucr = s->unst_ready[ib] ? kFTrue : 0;
    const fsibe_step_out &lb = s->last_b[ib];
    const double *F = lb.Force, *Fu = lb.Force_unst, *Fp = lb.Force_press;
    const double *IF = solid ? b->Inertial_Force : zero.data();
    ok = frec_write(f, {{s->bx.data() + o, B}, {s->by.data() + o, B}, {s->bvx.data() + o, B},
                        {s->bvy.data() + o, B}, {s->btheta.data() + o, B}, {s->btheta_dot.data() + o, B}}) &&
         frec_write(f, {{pn, B}, {po, B}, {tn, B}, {to, B}}) &&
         frec_write(f, {{F, 16}, {Fu, 16}, {Fp, 16}, {&Ek, 8}, {&Ep, 8}, {&Ee, 8}, {&x0, 8}, {&y0, 8}}) &&
         frec_write(f, {{&ab, 8}, {&imv, 4}, {&dun, 8}, {&dun, 8}, {&lun, 8}, {&lun, 8}}) &&
         frec_write(f, {{&ucr, 4}, {&dto, 8}, {solid ? b->oldold.data() : zero.data(), 6 * B}}) &&
         frec_write(f, {{solid ? b->ax.data() : zero.data(), B}, {solid ? b->ay.data() : zero.data(), B},
                        {IF, 16}});
  }
  ok = ok && frec_write(f, {{vm.data(), 8 * N}});
  ok = (fclose(f) == 0) && ok;
  if (!ok) return set_err(FSIBE_EARG, "backup: writing %s failed", path);
  return FSIBE_OK;
}

extern "C" int fsibe_backup_read(fsibe_sim *s, const char *path, int *ivideo, double *colorscale,
                                 int *colorscale_done) {
  if (!s || !path) return set_err(FSIBE_EARG, "null argument");
  const size_t N = (size_t)s->nx * s->ny;
  const int ns = s->ns;
  FILE *f = fopen(path, "rb");
  if (!f) return set_err(FSIBE_EARG, "backup: cannot open %s", path);
  std::vector<char> r;
  auto fail = [&](const char *what) {
    fclose(f);
    return set_err(FSIBE_EDOMAIN, "backup %s: %s", path, what);
  };
  double t = 0.0, chk = 0.0, dt1 = 0.0, cs = 0.0;
  int32_t in1 = 0, iv = 0, nsf = 0, cd = 0;
  std::vector<double> vk(2 * N), nl(2 * N), wv(2 * N), mk(3 * N), u(2 * N), vm(N);
  {
    if (!frec_read(f, r)) return fail("no time record");
    FCursor c(r);
    c.get(&t, 8);
    if (!c.done()) return fail("time record");
  }
  {
    if (!frec_read(f, r)) return fail("no checksum record");
    FCursor c(r);
    c.get(&chk, 8);
    if (!c.done() || chk != 123.0) return fail("checksum in backup file wrong (init_fields.f90:58-62)");
  }
  {
    if (!frec_read(f, r)) return fail("no field record");
    FCursor c(r);
    c.get(&in1, 4); c.get(&dt1, 8);
    c.get(vk.data(), 16 * N); c.get(nl.data(), 16 * N); c.get(wv.data(), 16 * N);
    c.get(mk.data(), 24 * N); c.get(&iv, 4); c.get(u.data(), 16 * N);
    if (!c.done()) return fail("field record does not match nx, ny");
    if (in1 != 0 && in1 != 1) return fail("time level index n1");
  }
  {
    if (!frec_read(f, r)) return fail("no ns record");
    FCursor c(r);
    c.get(&nsf, 4);
    if (!c.done() || nsf != s->ns) return fail("ns not the same in backup and params file");
  }
  {
    if (!frec_read(f, r)) return fail("no colorscale record");
    FCursor c(r);
    c.get(&cs, 8); c.get(&cd, 4);
    if (!c.done()) return fail("colorscale record");
  }
  const size_t B = 8 * (size_t)ns;
  // the six records of each beam (init_fields.f90:72-79)
  struct BeamRec {
    std::vector<double> arr, ld, sc, oo, acc, inf;
    double ab = 0.0, un[4] = {0, 0, 0, 0}, dto = 0.0;
    int32_t imv = 0, ucr = 0;
  };
  BeamRec br[2];
  for (int ib = 0; ib < s->nb; ++ib) {
    BeamRec &q = br[ib];
    q.arr.resize(6 * (size_t)ns); q.ld.resize(4 * (size_t)ns); q.sc.resize(11); q.oo.resize(6 * (size_t)ns);
    q.acc.resize(2 * (size_t)ns); q.inf.resize(2);
    bool ok = frec_read(f, r);
    if (ok) { FCursor c(r); c.get(q.arr.data(), 6 * B); ok = c.done(); }
    ok = ok && frec_read(f, r);
    if (ok) { FCursor c(r); c.get(q.ld.data(), 4 * B); ok = c.done(); }
    ok = ok && frec_read(f, r);
    if (ok) { FCursor c(r); c.get(q.sc.data(), 88); ok = c.done(); }
    ok = ok && frec_read(f, r);
    if (ok) { FCursor c(r); c.get(&q.ab, 8); c.get(&q.imv, 4); c.get(q.un, 32); ok = c.done(); }
    ok = ok && frec_read(f, r);
    if (ok) { FCursor c(r); c.get(&q.ucr, 4); c.get(&q.dto, 8); c.get(q.oo.data(), 6 * B); ok = c.done(); }
    ok = ok && frec_read(f, r);
    if (ok) { FCursor c(r); c.get(q.acc.data(), 2 * B); c.get(q.inf.data(), 16); ok = c.done(); }
    if (!ok) return fail("beam records");
  }
  {
    if (!frec_read(f, r)) return fail("no vor_mean record");
    FCursor c(r);
    c.get(vm.data(), 8 * N);
    if (!c.done()) return fail("vor_mean record");
  }
  fclose(f);

  // the state of the time loop after the backed-up step
  HIPCHK(hipStreamSynchronize(s->stream));
  const int n1 = in1, n0 = 1 - in1;
  s->n0 = n0;
  s->n1 = n1;
  CHK(upload_packed(s, vk.data() + n0 * N, s->W[n0]));
  CHK(upload_packed(s, vk.data() + n1 * N, s->W[n1]));
  CHK(upload_packed(s, nl.data() + n1 * N, s->NL[n1]));  // the next step's nlk(n1); nlk(n0) is rewritten
  HIPCHK(hipMemcpyAsync(s->u1, u.data(), sizeof(double) * N, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->u2, u.data() + N, sizeof(double) * N, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->vor_mean, vm.data(), sizeof(double) * N, hipMemcpyHostToDevice, s->stream));
  k_row_umax<<<s->ny, 256, 0, s->stream>>>(s->u1, s->u2, s->umax_part, s->nx, s->ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->time = t;
  s->dt1 = dt1;
  s->dt0 = dt1;
  s->it = 0;  // the loop counter restarts at 0 after a restart, as time_step.f90:174 sets it
  s->startup_done = true;
  s->have_beam = false;  // create_mask runs again from the beam (time_step.f90:84)
  s->res_pending = false;
  s->slot_fluid[0] = s->slot_fluid[1] = false;
  s->slot_forces[0] = s->slot_forces[1] = false;
  s->k1_ready = false;
  s->k2_ready = false;
  s->press_box_valid = false;
  s->beam_time = t - dt1;
  if (s->p.iFLUSI == 1 && s->nb > 0) CHK(ensure_solid(s));
  for (int ib = 0; ib < s->nb; ++ib) {
    const BeamRec &q = br[ib];
    const size_t o = (size_t)ib * ns;
    double *dst[6] = {s->bx.data() + o, s->by.data() + o, s->bvx.data() + o, s->bvy.data() + o,
                      s->btheta.data() + o, s->btheta_dot.data() + o};
    for (int k = 0; k < 6; ++k) memcpy(dst[k], q.arr.data() + (size_t)k * ns, B);
    memcpy(s->pbeam.data() + o, q.ld.data(), B);
    memcpy(s->tbeam.data() + o, q.ld.data() + 2 * (size_t)ns, B);
    fsibe_step_out &lb = s->last_b[ib];
    for (int k = 0; k < 2; ++k) {
      lb.Force[k] = q.sc[k];
      lb.Force_unst[k] = q.sc[2 + k];
      lb.Force_press[k] = q.sc[4 + k];
    }
    s->drag_unst_old[ib] = q.un[1];
    s->lift_unst_old[ib] = q.un[3];
    s->unst_ready[ib] = q.ucr != 0;
    if (s->p.iFLUSI == 1) {
      ibes::Beam &b = s->solid[ib]->b;
      std::vector<double> *ba[6] = {&b.x, &b.y, &b.vx, &b.vy, &b.theta, &b.theta_dot};
      for (int k = 0; k < 6; ++k) memcpy(ba[k]->data(), q.arr.data() + (size_t)k * ns, B);
      memcpy(b.pressure_new.data(), q.ld.data(), B);
      memcpy(b.pressure_old.data(), q.ld.data() + (size_t)ns, B);
      memcpy(b.tau_new.data(), q.ld.data() + 2 * (size_t)ns, B);
      memcpy(b.tau_old.data(), q.ld.data() + 3 * (size_t)ns, B);
      b.E_kinetic = q.sc[6]; b.E_pot = q.sc[7]; b.E_elastic = q.sc[8];
      b.dt_old = q.dto;
      memcpy(b.oldold.data(), q.oo.data(), 6 * B);
      memcpy(b.ax.data(), q.acc.data(), B);
      memcpy(b.ay.data(), q.acc.data() + ns, B);
      b.Inertial_Force[0] = q.inf[0];
      b.Inertial_Force[1] = q.inf[1];
      // not in the file: init_beam's .true. holds (init_beam.f90:107), so a
      // BDF2 run restarts with a Crank-Nicolson step, as the reference does
      b.StartupStep = true;
      s->solid[ib]->sv.pre_ok_ = false;
    }
  }
  sync_last(s);
  // the record's spectra and masks, for a backup written before the next step
  s->restored.resize(9 * N);
  std::copy(vk.begin(), vk.end(), s->restored.begin());
  std::copy(nl.begin(), nl.end(), s->restored.begin() + 2 * N);
  std::copy(wv.begin(), wv.end(), s->restored.begin() + 4 * N);
  std::copy(mk.begin(), mk.end(), s->restored.begin() + 6 * N);
  if (ivideo) *ivideo = iv;
  if (colorscale) *colorscale = cs;
  if (colorscale_done) *colorscale_done = (cd != 0) ? 1 : 0;
  return FSIBE_OK;
}

// write (10) nx, ny, press with press = cofitxy(vortk): the whole
// vortk(0:nx-1,0:ny-1,0:1) argument makes cofitxy read its first level, so
// the file holds the vorticity of level 0 (the older level after an odd
// number of steps)
extern "C" int fsibe_inicond_write(fsibe_sim *s, const char *path) {
  if (!s || !path) return set_err(FSIBE_EARG, "null argument");
  HIPCHK(hipStreamSynchronize(s->stream));
  const size_t N = (size_t)s->nx * s->ny;
  std::vector<double> v(N);
  double2 *outs[1] = {s->T[6]};
  const int fsel[1] = {0};
  CHK(cols_inv_fields(s, s->W[0], outs, fsel, 1));
  CHK(rows_c2r(s->nx, s->ny, s->px, s->twx, s->T[6], s->fb_sb, s->stream));
  HIPCHK(hipMemcpyAsync(v.data(), s->fb_sb, sizeof(double) * N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->press_box_valid = false;
  FILE *f = fopen(path, "wb");
  if (!f) return set_err(FSIBE_EARG, "inicond: cannot open %s for writing", path);
  const int32_t nx = s->nx, ny = s->ny;
  bool ok = frec_write(f, {{&nx, 4}, {&ny, 4}, {v.data(), 8 * N}});
  ok = (fclose(f) == 0) && ok;
  if (!ok) return set_err(FSIBE_EARG, "inicond: writing %s failed", path);
  return FSIBE_OK;
}

// read (10) nx_file, ny_file, vort; vortk(:,:,0) = coftxy(vort); the startup
// step computes u from it (EvolveFluidExplicit)
extern "C" int fsibe_inicond_read(fsibe_sim *s, const char *path) {
  if (!s || !path) return set_err(FSIBE_EARG, "null argument");
  if (s->startup_done || s->n0 != 0) return set_err(FSIBE_ESTATE, "inicond: the simulation has already stepped");
  const size_t N = (size_t)s->nx * s->ny;
  FILE *f = fopen(path, "rb");
  if (!f) return set_err(FSIBE_EARG, "inicond: cannot open %s", path);
  std::vector<char> r;
  const bool got = frec_read(f, r);
  fclose(f);
  if (!got) return set_err(FSIBE_EDOMAIN, "inicond %s: no record", path);
  FCursor c(r);
  int32_t nx = 0, ny = 0;
  c.get(&nx, 4); c.get(&ny, 4);
  if (!c.ok || nx != s->nx || ny != s->ny)
    return set_err(FSIBE_EDOMAIN, "inicond field resolution does not match present resolution (init_fields.f90:118-126)");
  std::vector<double> v(N);
  c.get(v.data(), 8 * N);
  if (!c.done()) return set_err(FSIBE_EDOMAIN, "inicond %s: record length", path);
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpyAsync(s->fb_sa, v.data(), sizeof(double) * N, hipMemcpyHostToDevice, s->stream));
  CHK(rows_r2c(s->nx, s->ny, s->px, s->twx, s->fb_sa, s->T[6], s->stream));
  const double norm = 1.0 / ((double)s->nx * (double)s->ny);
  CHK(with_cols<1>(s->py, s->H, pick_tx(s->H, s->ny, 1, 1), [&](auto F, int nt, int TX) -> int {
    const size_t lds = sizeof(double2) * TX * padlen(s->ny);
    KLAUNCH(k_cols_fwd_pack<decltype(F)>, s->H / TX, nt, lds, s->stream, s->T[6], s->fb_p, F, s->twy, s->H, norm);
    return FSIBE_OK;
  }));
  k_packed_to_half<<<grid1d((size_t)s->H * (s->ny / 2)), 256, 0, s->stream>>>(s->fb_p, s->W[0], s->nx, s->ny);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->press_box_valid = false;
  return FSIBE_OK;
}

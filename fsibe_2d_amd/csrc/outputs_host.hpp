// outputs_host.hpp -- field outputs of the reference's save_fields path;
// part of fsibe_hip.hip's translation unit (included at its end).
//
//   fsibe_pressure_snapshot  ComputePressureSnapshot (NavierStokes.f90:598-665)
//   fsibe_save_field         SaveField (FieldExport.f90:612-648)

// ---- ComputePressureSnapshot -------------------------------------------------
// Not part of the time step: the static pressure of a stored vorticity level
// for a field file, computed like the reference with the packed-layout
// spectral operators (the same device operators as fsibe_dev_*):
//   psi = poisson(vortk); u = (cofitxy(cofdy psi), -cofitxy(cofdx psi)) + U Mean_u(time)
//   penal = mask (u - u_s);  vort = cofitxy(vortk)
//   rhs = cofdx(coftxy(penal_1)) + cofdy(coftxy(penal_2))
//         + cofdy(coftxy(vort u_1)) + cofdx(coftxy(-vort u_2))
//   press = cofitxy(poisson(rhs)) - (u_1^2 + u_2^2) / 2

// u1 <- u1 + Ux, u2 <- -u2 + Uy (cal_velocity, NavierStokes.f90:206-209)
__global__ void k_snap_vel(double *__restrict__ u1, double *__restrict__ u2, double Ux, double Uy, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    u1[i] = u1[i] + Ux;
    u2[i] = -u2[i] + Uy;
  }
}

// penal = mask (u - u_s) on the mask box, 0 elsewhere (:624-625); wdx =
// -vort u2, wdy = +vort u1 (:643-644)
__global__ void k_snap_terms(const double *__restrict__ u1, const double *__restrict__ u2,
                             const double *__restrict__ vort, MaskBox mb, int nx, long long n,
                             double *__restrict__ p1, double *__restrict__ p2, double *__restrict__ wdx,
                             double *__restrict__ wdy) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int y = (int)(i / nx), x = (int)(i - (long long)y * nx);
    double a = 0.0, b = 0.0;
    if (mb.x1 >= mb.x0 && mb.inside(x, y)) {
      const size_t q = (size_t)(y - mb.y0) * mb.w + (x - mb.x0);
      a = mb.m[q] * (u1[i] - mb.vx[q]);
      b = mb.m[q] * (u2[i] - mb.vy[q]);
    }
    p1[i] = a;
    p2[i] = b;
    wdx[i] = -vort[i] * u2[i];
    wdy[i] = +vort[i] * u1[i];
  }
}

// rhs = penal_1 + penal_2 + work_dy + work_dx, left to right (:649-650)
__global__ void k_snap_sum(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
                           const double *__restrict__ d, double *__restrict__ o, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    o[i] = ((a[i] + b[i]) + c[i]) + d[i];
}

// press = work1 - 0.5 (u1^2 + u2^2) (:663)
__global__ void k_snap_press(const double *__restrict__ w, const double *__restrict__ u1,
                             const double *__restrict__ u2, double *__restrict__ o, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    o[i] = w[i] - 0.5 * (u1[i] * u1[i] + u2[i] * u2[i]);
}

extern "C" int fsibe_pressure_snapshot(fsibe_sim *s, double time, int level, double *press) {
  if (!s || !press) return set_err(FSIBE_EARG, "null argument");
  if (level != 0 && level != 1) return set_err(FSIBE_EARG, "level must be 0 or 1");
  HIPCHK(hipStreamSynchronize(s->stream));
  const size_t N = (size_t)s->nx * s->ny;
  if (!s->spec) CHK(fsibe_spectral_create(s->nx, s->ny, s->p.xl, s->p.yl, s->p.nu, &s->spec));
  if (!s->snap) HIPCHK(hipMalloc(&s->snap, sizeof(double) * 9 * N));
  double *vk = s->snap, *tmp = vk + N, *u1 = tmp + N, *u2 = u1 + N, *vort = u2 + N, *p1 = vort + N, *p2 = p1 + N,
         *wdx = p2 + N, *wdy = wdx + N;
  fsibe_spectral *c = s->spec;
  hipStream_t st = s->stream;
  const int g = grid1d(N);
  k_half_to_packed<<<grid1d((size_t)s->H * (s->ny / 2)), 256, 0, st>>>(s->W[level], vk, s->nx, s->ny);
  HIPCHK(hipGetLastError());
  // cal_velocity (NavierStokes.f90:181-212)
  double mx, my;
  mean_u_host(s->fp, time, mx, my);
  if (s->p.iMeanVelocity == 11) {
    DevScalars d;
    HIPCHK(hipMemcpyAsync(&d, s->dscal, sizeof(DevScalars), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    mx = d.u_mean[0];
    my = d.u_mean[1];
  }
  CHK(dev_op(c, 5, vk, wdx, st));  // stream function (poisson)
  CHK(dev_op(c, 0, wdx, tmp, st));
  CHK(dev_cofitxy(c, tmp, u2, st));  // dpsi/dx = -u_y
  CHK(dev_op(c, 1, wdx, tmp, st));
  CHK(dev_cofitxy(c, tmp, u1, st));  // dpsi/dy = u_x
  k_snap_vel<<<g, 256, 0, st>>>(u1, u2, s->p.U_mean_true * mx, s->p.U_mean_true * my, (long long)N);
  HIPCHK(hipGetLastError());
  CHK(dev_cofitxy(c, vk, vort, st));
  k_snap_terms<<<g, 256, 0, st>>>(u1, u2, vort, mask_box(s), s->nx, (long long)N, p1, p2, wdx, wdy);
  HIPCHK(hipGetLastError());
  // divergences in Fourier space, each into its own array (:630-648)
  CHK(dev_coftxy(c, p1, tmp, st));
  CHK(dev_op(c, 0, tmp, p1, st));
  CHK(dev_coftxy(c, p2, tmp, st));
  CHK(dev_op(c, 1, tmp, p2, st));
  CHK(dev_coftxy(c, wdx, tmp, st));
  CHK(dev_op(c, 0, tmp, wdx, st));
  CHK(dev_coftxy(c, wdy, tmp, st));
  CHK(dev_op(c, 1, tmp, wdy, st));
  k_snap_sum<<<g, 256, 0, st>>>(p1, p2, wdy, wdx, vort, (long long)N);
  HIPCHK(hipGetLastError());
  CHK(dev_op(c, 5, vort, tmp, st));  // poisson
  CHK(dev_cofitxy(c, tmp, vk, st));
  k_snap_press<<<g, 256, 0, st>>>(vk, u1, u2, tmp, (long long)N);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(press, tmp, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return FSIBE_OK;
}

// ---- SaveField ---------------------------------------------------------------
// The reference's field file: the PARAMS header (WriteHeaderToFile,
// BeamIO.f90:364-376: each line as a CHARACTER(256) record, blank-padded),
// "%format1", nx and ny as I5.5, xl and yl as ES11.4, the field type, then
// one value per line from the top row down (field(ix, ny1-iy), every
// nsave-th row and column) as ES16.9 for the type "precision", else ES11.4.

// Fortran ESw.d output of v (w = d + 7, two exponent digits, or three with
// the E dropped, as the standard prescribes for |exponent| > 99)
static void fortran_es(char *out, size_t cap, double v, int d) {
  const int w0 = d + 7;
  if (!std::isfinite(v)) {
    snprintf(out, cap, "%*s", w0, std::isnan(v) ? "NaN" : (v > 0 ? "Infinity" : "-Infinity"));
    return;
  }
  char buf[64];
  snprintf(buf, sizeof buf, "%.*E", d, v);
  // buf = [-]D.DDDDE[+-]XX[X]
  char *e = strchr(buf, 'E');
  const int w = d + 7;
  char mant[40], exps[8];
  const size_t ml = (size_t)(e - buf);
  memcpy(mant, buf, ml);
  mant[ml] = '\0';
  const char sign = e[1];
  const char *digs = e + 2;
  while (strlen(digs) > 2 && digs[0] == '0') ++digs;
  if (strlen(digs) <= 2) snprintf(exps, sizeof exps, "E%c%s", sign, digs);
  else snprintf(exps, sizeof exps, "%c%s", sign, digs);
  char body[64];
  snprintf(body, sizeof body, "%s%s", mant, exps);
  const int len = (int)strlen(body);
  if (len > w) {  // does not fit: the standard's asterisks
    memset(out, '*', (size_t)w);
    out[w] = '\0';
    return;
  }
  snprintf(out, cap, "%*s", w, body);
}

extern "C" int fsibe_save_field(const char *path, const double *field, int nx, int ny, int nsave, double xl,
                                double yl, const char *fieldtype, const char *header) {
  if (!path || !field || !fieldtype || nx < 1 || ny < 1 || nsave < 1) return set_err(FSIBE_EARG, "bad argument");
  FILE *f = fopen(path, "w");
  if (!f) return set_err(FSIBE_EARG, "SaveField: cannot open %s", path);
  if (header) {
    const char *p = header;
    while (*p) {
      const char *e = strchr(p, '\n');
      const size_t n = e ? (size_t)(e - p) : strlen(p);
      char line[257];
      const size_t k = std::min<size_t>(n, 256);
      memcpy(line, p, k);
      memset(line + k, ' ', 256 - k);
      line[256] = '\0';
      fprintf(f, "%s\n", line);
      p += n + (e ? 1 : 0);
    }
  }
  char a[32], b[32];
  fortran_es(a, sizeof a, xl, 4);
  fortran_es(b, sizeof b, yl, 4);
  fprintf(f, "%%format1\n%05d   %% nx\n%05d   %% ny\n%s   %% xl, domain size in x-direction\n"
             "%s   %% yl, domain size in y-direction\n"
             "%% The next entry can be used to identify the type of the field: \n%% %s\n",
          nx, ny, a, b, fieldtype);
  const int d = strcmp(fieldtype, "precision") == 0 ? 9 : 4;
  char v[40];
  for (int iy = 0; iy <= ny - 1; iy += nsave) {
    const int row = ny - 1 - iy;  // field(ix, ny1-iy), 1-based
    for (int ix = 0; ix < nx; ix += nsave) {
      fortran_es(v, sizeof v, field[(size_t)row * nx + ix], d);
      fprintf(f, "%s\n", v);
    }
  }
  const bool ok = fclose(f) == 0;
  if (!ok) return set_err(FSIBE_EARG, "SaveField: writing %s failed", path);
  return FSIBE_OK;
}

// outputs_host.hpp -- field outputs of the reference's save_fields path;
// part of fsibe_hip.hip's translation unit (included at its end).
//
//   fsibe_pressure_snapshot  ComputePressureSnapshot (NavierStokes.f90:598-665)
//   fsibe_save_field         SaveField (FieldExport.f90:612-648)
//   fsibe_save_gif           SaveGIF (FieldExport.f90:305-384) with the GIF89a
//                            writer of gif_util.f90

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

// ---- SaveGIF -----------------------------------------------------------------
// A field as a 256-colour GIF: each value is mapped to a palette index,
// 1 + nint(255 (f - min) / (max - min)) clamped to 1 .. 256 (minus one),
// row ny-1 of the field at the top of the image (SaveGIF, FieldExport.f90:
// 336-380), and the index raster is LZW-coded as gif_util.f90 does it.

// Fortran nint: the nearest integer, halves away from zero
static long fnint_l(double v) { return std::lround(v); }

// the computed palettes of FargePalette (FieldExport.f90:1653-1711): 11
// (Matlab "spring"), 13 (Matlab's standard map, SaveGIF's default) and 14
// (13 in 16 flat bands); rgb[3][256].  The tabulated palettes 1-10, and 12
// (whose first and last entries FargePalette leaves unset), are not built.
static bool gif_palette(int ip, int rgb[3][256]) {
  int p13[3][256];
  for (int i = 1; i <= 256; ++i) {
    const double t = (double)(i - 1) / 255.0;
    int r, g, b;
    if (t <= 0.20) {
      r = 0; g = 0; b = 143 + (int)fnint_l(t * (255.0 - 143.0) * 5.0);
    } else if (t > 0.20 && t <= 0.40) {
      r = 0; g = (int)fnint_l((t - 0.20) * 255.0 * 5.0); b = 255;
    } else if (t > 0.40 && t <= 0.60) {
      r = (int)fnint_l((t - 0.40) * 255.0 * 5.0); g = 255; b = 255 - (int)fnint_l((t - 0.40) * 255.0 * 5.0);
    } else if (t > 0.60 && t <= 0.80) {
      r = 255; g = 255 - (int)fnint_l((t - 0.60) * 255.0 * 5.0); b = 0;
    } else {
      r = 255 - (int)fnint_l((t - 0.80) * (255.0 - 128.0) * 5.0); g = 0; b = 0;
    }
    p13[0][i - 1] = r; p13[1][i - 1] = g; p13[2][i - 1] = b;
    if (ip == 11) {
      rgb[0][i - 1] = (int)fnint_l(t * 255.0);
      rgb[1][i - 1] = (int)fnint_l(255.0 - t * 255.0);
      rgb[2][i - 1] = 255;
    }
  }
  if (ip == 11) return true;
  if (ip == 13) {
    memcpy(rgb, p13, sizeof p13);
    return true;
  }
  if (ip == 14) {  // 16 bands of nint(255/16) = 16 entries, each the colour of its first
    const int n = 16, sector = (int)fnint_l(255.0 / (double)n);
    for (int j = 0; j < n; ++j)
      for (int i = j * sector + 1; i <= (j + 1) * sector; ++i)
        for (int c = 0; c < 3; ++c) rgb[c][i - 1] = p13[c][j * sector];
    return true;
  }
  return false;
}

// The LZW raster of gif_util.f90 (giflzw, slicewrite, flushbuffer): 8-bit
// pixels, codes from 9 bits up; a code's width grows when the next code to
// be assigned passes the current maximum; at 4096 codes a clear code
// restarts the table; data sub-blocks of at most 255 bytes.
class GifLzw {
 public:
  explicit GifLzw(std::vector<unsigned char> &out) : out_(out) {}
  void code(int c) {
    accum_ += (unsigned long long)c << skip_;
    skip_ += slen_;
    buf_[ibuf_] = (unsigned char)(accum_ & 255u);
    while (skip_ >= 8) {
      ++ibuf_;
      accum_ >>= 8;
      skip_ -= 8;
      buf_[ibuf_] = (unsigned char)(accum_ & 255u);
    }
    if (ibuf_ > 255) flush();
  }
  // writes up to 255 complete bytes, or at the end the partial last byte too
  void flush() {
    int bl;
    if (ibuf_ > 255) bl = 255;
    else if (skip_ != 0) bl = ibuf_;
    else if (ibuf_ > 1) bl = ibuf_ - 1;
    else return;
    out_.push_back((unsigned char)bl);
    out_.insert(out_.end(), buf_ + 1, buf_ + 1 + bl);
    memmove(buf_ + 1, buf_ + 1 + bl, (size_t)(ibuf_ - bl));
    ibuf_ -= bl;
  }
  int slen_ = 9;

 private:
  std::vector<unsigned char> &out_;
  unsigned char buf_[260] = {0};  // buf_[1 .. ibuf_], buf_[ibuf_] the partial byte
  int ibuf_ = 1, skip_ = 0;
  unsigned long long accum_ = 0;
};

static void gif_raster(std::vector<unsigned char> &out, const std::vector<unsigned char> &px) {
  const int blen = 8, maxbase = 255, CC = 256, EOI = 257, maxcode = 4095;
  out.push_back((unsigned char)blen);
  GifLzw w(out);
  // the table: code of the sequence [P]K, looked up by (P << 8) | K
  std::vector<int> table((size_t)(maxcode + 1) << 8, -1);
  std::vector<int> used;  // entries to clear on a reset
  int ncod = 0, curmaxcode = 0;
  auto init = [&]() {
    for (int e : used) table[(size_t)e] = -1;
    used.clear();
    ncod = CC + 2;
    w.slen_ = blen + 1;
    curmaxcode = (1 << w.slen_) - 1;
  };
  init();
  w.code(CC);
  int P = -1;
  for (size_t q = 0; q < px.size(); ++q) {
    const int K = px[q] & maxbase;
    if (q == 0) {
      P = K;
      continue;
    }
    const int e = (P << 8) | K;
    if (table[(size_t)e] >= 0) {
      P = table[(size_t)e];
      continue;
    }
    w.code(P);
    if (ncod > maxcode) {
      w.code(CC);
      init();
    } else {
      if (ncod > curmaxcode) {
        w.slen_ += 1;
        curmaxcode = curmaxcode * 2 + 1;
      }
      table[(size_t)e] = ncod;
      used.push_back(e);
      ++ncod;
    }
    P = K;
  }
  w.code(P);
  w.code(EOI);
  w.flush();
}

extern "C" int fsibe_save_gif(const char *path, const double *field, int nx, int ny, int palette, int have_range,
                              double color_min, double color_max) {
  if (!path || !field || nx < 1 || ny < 1 || nx > 65535 || ny > 65535) return set_err(FSIBE_EARG, "bad argument");
  int rgb[3][256];
  if (!gif_palette(palette, rgb))
    return set_err(FSIBE_EUNSUPPORTED, "SaveGIF: palette %d (only the computed palettes 11, 13 and 14)", palette);
  double minf = color_min, maxf = color_max;
  if (!have_range) {  // minval / maxval of the field
    minf = maxf = field[0];
    for (size_t q = 1; q < (size_t)nx * ny; ++q) {
      minf = std::min(minf, field[q]);
      maxf = std::max(maxf, field[q]);
    }
  }
  // Pixel(i, j) for field(i, ny+1-j): raster rows from the field's last row
  std::vector<unsigned char> px((size_t)nx * ny);
  for (int j = 0; j < ny; ++j) {
    const double *row = field + (size_t)(ny - 1 - j) * nx;
    for (int i = 0; i < nx; ++i) {
      const double v = 255.0 * (row[i] - minf) / (maxf - minf);
      long it = std::isnan(v) ? 1 : 1 + fnint_l(v);
      if (it < 1) it = 1;
      if (it > 256) it = 256;
      px[(size_t)j * nx + i] = (unsigned char)(it - 1);
    }
  }
  std::vector<unsigned char> g;
  const char hdr[] = "GIF89a";
  g.insert(g.end(), hdr, hdr + 6);
  auto u16 = [&](int v) { g.push_back((unsigned char)(v & 255)); g.push_back((unsigned char)((v >> 8) & 255)); };
  u16(nx);
  u16(ny);
  g.push_back((unsigned char)(128 + 7 * 16 + 7));  // global colour table of 2^8 entries
  g.push_back(0);                                  // background
  g.push_back(0);                                  // aspect
  for (int i = 0; i < 256; ++i)
    for (int c = 0; c < 3; ++c) g.push_back((unsigned char)rgb[c][i]);
  g.push_back(',');
  u16(0);
  u16(0);
  u16(nx);
  u16(ny);
  g.push_back(7);  // no local table, not interlaced, 8-bit pixels
  gif_raster(g, px);
  g.push_back(0);
  g.push_back(';');
  std::string name = std::string(path) + ".gif";  // SaveGIF appends the extension
  FILE *f = fopen(name.c_str(), "wb");
  if (!f) return set_err(FSIBE_EARG, "SaveGIF: cannot open %s", name.c_str());
  const bool ok = fwrite(g.data(), 1, g.size(), f) == g.size();
  if (fclose(f) != 0 || !ok) return set_err(FSIBE_EARG, "SaveGIF: writing %s failed", name.c_str());
  return FSIBE_OK;
}

// Synthetic projection inputs for the five benchmark configurations.
//
// This is harness code, not part of the drop-in engine: it produces the
// arrays a SIMCOFlow host would hand to the projection (cut-cell geometry,
// two-phase level set / VOF, momentum d-coefficients and a seeded
// intermediate velocity) from analytic shapes.  The semantics it follows:
//   * cell classification by fluid fraction with epsi, and the vof/vofL
//     normalisation of full cells           -- CutCell.f90:329-365
//   * "all faces < 1e-2" demotion to solid, zeroing of solid-cell faces and
//     of cut-cell faces that touch a solid   -- CutCell.f90:543-596
//   * momentum-exchange (MoEx) U/V/W cells   -- CutCell.f90:673-749
//   * u* = 0 on no-slip wall-normal planes   -- PredictUVW.f90:3796-3894
//     and in solid U/V/W cells               -- PredictUVW.f90:3896-3916
//   * staggered centre distances             -- Mesh.f90:160-247
//   * Dp = dt on every face: DiffusiveFlux zeroes the viscous fluxes
//     (PredictUVW.f90:3672-3673) so 1/(aP - sum a_nb) = dt/(vof V) and the
//     vof V scaling (PredictUVW.f90:1230-1247) gives dt.
// Faces whose U/V/W cell is solid or MoEx are closed (aperture 0), so the
// pressure operator sees exactly the faces the correction keeps open.
//
// Layouts: "halo" arrays are Fortran column-major (0:I+1, 0:J+1, 0:K+1);
// "interior" arrays are (1:I, 1:J, 1:K); 1-D spacings have length I, J, K.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEpsi = 1.0e-3;   // CutCell.f90:11

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based uniform in [0,1): stream and index select the draw.
inline double u01(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t key = mix64(seed ^ (stream << 56));
  return (double)(mix64(key ^ index) >> 11) * (1.0 / 9007199254740992.0);
}

struct Box { double lo[3], hi[3]; };

struct Setup {
  int cfg = 0;
  int I = 0, J = 0, K = 0;
  uint64_t seed = 0;
  double L[3] = {1, 1, 1};
  // solid
  int solid = 0;        // 0 none, 1 sphere, 2 x-cylinder, 3 boxes
  double sc[3] = {0, 0, 0}, sr = 0;
  Box boxes[3];
  int nbox = 0;
  // liquid surface eta(x,y)
  int surf = 0;
  int km[8][2];
  double kph[8];
  double wave_phase = 0;
  double eta_lip = 1.0;  // bound on |d eta/dx| + |d eta/dy|
  double ustar_amp = 1.0;
  // z stretching
  int zstretch = 0;
};

double eta(const Setup& s, double x, double y) {
  switch (s.surf) {
    case 0: return 0.5 + 0.05 * std::sin(2 * kPi * x) * std::cos(2 * kPi * y);
    case 1: {
      double acc = 0;
      for (int m = 0; m < 8; ++m)
        acc += std::cos(2 * kPi * (s.km[m][0] * x + s.km[m][1] * y) + s.kph[m]);
      return 0.6 + 0.02 * acc / 8.0;
    }
    case 2:
      return 0.5 + 0.1 * std::tanh(20.0 * (x - 0.5)) * (1.0 + 0.3 * std::sin(4 * kPi * y));
    case 3:
      return 0.5 + 0.03 * std::cos(2 * kPi * x / 1.0 + s.wave_phase);
  }
  return 0.5;
}

// Solid level set: > 0 in fluid.  Returns also a Lipschitz-type bound so a
// whole cell can be classified from its centre.
double phi_solid(const Setup& s, double x, double y, double z) {
  switch (s.solid) {
    case 1: {
      double dx = x - s.sc[0], dy = y - s.sc[1], dz = z - s.sc[2];
      return std::sqrt(dx * dx + dy * dy + dz * dz) - s.sr;
    }
    case 2: {
      double dy = y - s.sc[1], dz = z - s.sc[2];
      return std::sqrt(dy * dy + dz * dz) - s.sr;
    }
    case 3: {
      double best = 1e300;
      for (int b = 0; b < s.nbox; ++b) {
        const Box& B = s.boxes[b];
        double c0 = 0.5 * (B.lo[0] + B.hi[0]), c1 = 0.5 * (B.lo[1] + B.hi[1]),
               c2 = 0.5 * (B.lo[2] + B.hi[2]);
        double h0 = 0.5 * (B.hi[0] - B.lo[0]), h1 = 0.5 * (B.hi[1] - B.lo[1]),
               h2 = 0.5 * (B.hi[2] - B.lo[2]);
        double d = std::max({std::fabs(x - c0) - h0, std::fabs(y - c1) - h1,
                             std::fabs(z - c2) - h2});
        best = std::min(best, d);
      }
      return best;
    }
  }
  return 1e300;
}

struct Grid1 {
  std::vector<double> f;   // faces 0..n
  std::vector<double> c;   // centres 0..n+1 (ghosts extrapolated)
  std::vector<double> d;   // widths 1..n stored at [1..n], ghosts mirror
  std::vector<double> dc;  // staggered centre distance 1..n
  int n = 0;
};

void build_uniform(Grid1& g, int n, double L) {
  g.n = n;
  g.f.resize(n + 1);
  for (int i = 0; i <= n; ++i) g.f[i] = L * (double)i / (double)n;
}

// Cells shrink geometrically (ratio r per cell) toward z = zi from both
// sides; each side is scaled so that it fills [0, zi] and [zi, L] exactly.
void build_stretched(Grid1& g, int n, double L, double zi, double r) {
  g.n = n;
  int nb = (int)std::lround(n * zi / L);
  nb = std::min(std::max(nb, 1), n - 1);
  int na = n - nb;
  auto side = [&](int m, double len, std::vector<double>& out) {
    out.resize(m);
    double sum = 0;
    for (int q = 0; q < m; ++q) { out[q] = std::pow(r, q); sum += out[q]; }
    for (int q = 0; q < m; ++q) out[q] *= len / sum;
  };
  std::vector<double> below, above;
  side(nb, zi, below);      // below[0] touches the interface
  side(na, L - zi, above);  // above[0] touches the interface
  g.f.resize(n + 1);
  g.f[0] = 0;
  for (int q = 0; q < nb; ++q) g.f[q + 1] = g.f[q] + below[nb - 1 - q];
  g.f[nb] = zi;
  for (int q = 0; q < na; ++q) g.f[nb + q + 1] = g.f[nb + q] + above[q];
  g.f[n] = L;
}

void finish_grid(Grid1& g) {
  int n = g.n;
  g.d.assign(n + 2, 0);
  g.c.assign(n + 2, 0);
  g.dc.assign(n + 2, 0);
  for (int i = 1; i <= n; ++i) {
    g.d[i] = g.f[i] - g.f[i - 1];
    g.c[i] = 0.5 * (g.f[i] + g.f[i - 1]);
  }
  g.d[0] = g.d[1];
  g.d[n + 1] = g.d[n];
  g.c[0] = g.f[0] - 0.5 * g.d[0];
  g.c[n + 1] = g.f[n] + 0.5 * g.d[n + 1];
  for (int i = 1; i <= n; ++i) g.dc[i] = (i < n) ? (g.c[i + 1] - g.c[i]) : g.d[i];
}

// cell extent in one direction, valid for ghost cells too
inline void cell_span(const Grid1& g, int i, double& a, double& b) {
  if (i == 0) { b = g.f[0]; a = b - g.d[0]; }
  else if (i == g.n + 1) { a = g.f[g.n]; b = a + g.d[g.n + 1]; }
  else { a = g.f[i - 1]; b = g.f[i]; }
}

constexpr int NS = 8;  // sub-samples per direction

// fluid fraction and liquid-and-fluid fraction of an axis-aligned box
void box_fractions(const Setup& s, const double lo[3], const double hi[3],
                   double& fluid, double& liquid) {
  double h[3] = {0.5 * (hi[0] - lo[0]), 0.5 * (hi[1] - lo[1]), 0.5 * (hi[2] - lo[2])};
  double c[3] = {lo[0] + h[0], lo[1] + h[1], lo[2] + h[2]};
  double ps = phi_solid(s, c[0], c[1], c[2]);
  double rad = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  bool all_fluid = ps > rad, all_solid = ps < -rad;
  double pl = c[2] - eta(s, c[0], c[1]);
  double lrad = h[2] + s.eta_lip * std::max(h[0], h[1]);
  bool all_liq = pl < -lrad, all_gas = pl > lrad;
  if (all_solid) { fluid = 0; liquid = 0; return; }
  if (all_fluid && (all_liq || all_gas)) { fluid = 1; liquid = all_liq ? 1 : 0; return; }
  int nf = 0, nl = 0;
  for (int a = 0; a < NS; ++a)
    for (int b = 0; b < NS; ++b)
      for (int q = 0; q < NS; ++q) {
        double x = lo[0] + (a + 0.5) * (hi[0] - lo[0]) / NS;
        double y = lo[1] + (b + 0.5) * (hi[1] - lo[1]) / NS;
        double z = lo[2] + (q + 0.5) * (hi[2] - lo[2]) / NS;
        bool fl = all_fluid ? true : (phi_solid(s, x, y, z) > 0);
        if (fl) {
          ++nf;
          bool lq = all_liq ? true : (all_gas ? false : (z - eta(s, x, y) < 0));
          if (lq) ++nl;
        }
      }
  fluid = (double)nf / (NS * NS * NS);
  liquid = (double)nl / (NS * NS * NS);
}

// fluid fraction of an axis-aligned face (dir = normal axis)
double face_fraction(const Setup& s, int dir, double pos, const double lo[3], const double hi[3]) {
  double p[3], h[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) { p[d] = 0.5 * (lo[d] + hi[d]); if (d != dir) h[d] = 0.5 * (hi[d] - lo[d]); }
  p[dir] = pos;
  double ps = phi_solid(s, p[0], p[1], p[2]);
  double rad = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  if (ps > rad) return 1.0;
  if (ps < -rad) return 0.0;
  int t0 = (dir + 1) % 3, t1 = (dir + 2) % 3, nf = 0;
  for (int a = 0; a < NS; ++a)
    for (int b = 0; b < NS; ++b) {
      double q[3];
      q[dir] = pos;
      q[t0] = lo[t0] + (a + 0.5) * (hi[t0] - lo[t0]) / NS;
      q[t1] = lo[t1] + (b + 0.5) * (hi[t1] - lo[t1]) / NS;
      if (phi_solid(s, q[0], q[1], q[2]) > 0) ++nf;
    }
  return (double)nf / (NS * NS);
}

Setup make_setup(int cfg, int I, int J, int K, uint64_t seed) {
  Setup s;
  s.cfg = cfg;
  s.seed = seed;
  int dI[5] = {64, 128, 256, 128, 512}, dJ[5] = {64, 128, 256, 128, 256}, dK[5] = {64, 128, 256, 128, 256};
  s.I = I > 0 ? I : dI[cfg];
  s.J = J > 0 ? J : dJ[cfg];
  s.K = K > 0 ? K : dK[cfg];
  switch (cfg) {
    case 0:
    case 3:
      s.surf = 0;
      s.eta_lip = 0.05 * 2 * kPi;
      s.solid = 1; s.sc[0] = 0.5; s.sc[1] = 0.5; s.sc[2] = 0.3; s.sr = 0.15;
      break;
    case 1:
      s.surf = 1;
      for (int m = 0; m < 8; ++m) {
        s.km[m][0] = 1 + (int)(4.0 * u01(seed, 10, 3 * m));
        s.km[m][1] = 1 + (int)(4.0 * u01(seed, 10, 3 * m + 1));
        s.kph[m] = 2 * kPi * u01(seed, 10, 3 * m + 2);
      }
      s.eta_lip = 0.02 * 2 * kPi * 4.0;
      s.solid = 2; s.sc[1] = 0.5; s.sc[2] = 0.35; s.sr = 0.1;
      s.zstretch = 1;
      break;
    case 2:
      s.surf = 2;
      s.eta_lip = 0.1 * 20.0 * 1.3 + 0.1 * 0.3 * 4 * kPi;
      s.solid = 0;
      break;
    case 4: {
      s.L[0] = 2.0; s.L[1] = 1.0; s.L[2] = 1.0;  // dx = 1/256 at 512x256x256
      s.surf = 3;
      s.wave_phase = 2 * kPi * u01(seed, 10, 0);
      s.eta_lip = 0.03 * 2 * kPi;
      s.solid = 3; s.nbox = 3;
      for (int b = 0; b < 3; ++b) {
        double side[3];
        for (int d = 0; d < 3; ++d) side[d] = 0.05 + 0.10 * u01(seed, 11, 10 * b + d);
        double cx = 0.1 + side[0] / 2 + (1.8 - side[0]) * u01(seed, 11, 10 * b + 3);
        double cy = 0.05 + side[1] / 2 + (0.9 - side[1]) * u01(seed, 11, 10 * b + 4);
        // fully submerged: top below min(eta) - 0.05, bottom above 0.05
        double top_max = 0.47 - 0.05, bot_min = 0.05;
        double cz = bot_min + side[2] / 2 + (top_max - bot_min - side[2]) * u01(seed, 11, 10 * b + 5);
        s.boxes[b].lo[0] = cx - side[0] / 2; s.boxes[b].hi[0] = cx + side[0] / 2;
        s.boxes[b].lo[1] = cy - side[1] / 2; s.boxes[b].hi[1] = cy + side[1] / 2;
        s.boxes[b].lo[2] = cz - side[2] / 2; s.boxes[b].hi[2] = cz + side[2] / 2;
      }
      s.ustar_amp = 0.5;
      break;
    }
  }
  return s;
}

}  // namespace

extern "C" {

// Output buffers; the caller owns them.  Halo arrays have (I+2)(J+2)(K+2)
// doubles, interior int arrays I*J*K, 1-D arrays I, J or K doubles.
typedef struct sg_fields {
  double *dxp, *dyp, *dzp, *dxu, *dyv, *dzw;   // 1-D spacings
  double *xc, *yc, *zc;                         // 1-D cell centres
  double *u_star, *v_star, *w_star;
  double *dp_u, *dp_v, *dp_w;
  double *ee_area, *ne_area, *te_area;
  double *vof, *vof_old, *vof_l, *phi_l;
  int32_t *cell_type_p, *cell_type_u, *cell_type_v, *cell_type_w;
  int32_t *moex_u, *moex_v, *moex_w;
} sg_fields;

// Default sizes of a configuration (0..4).
int sg_config_dims(int cfg, int32_t* dims) {
  if (cfg < 0 || cfg > 4) return -1;
  Setup s = make_setup(cfg, 0, 0, 0, 0);
  dims[0] = s.I; dims[1] = s.J; dims[2] = s.K;
  return 0;
}

// Domain lengths of a configuration.
int sg_config_lengths(int cfg, double* L) {
  if (cfg < 0 || cfg > 4) return -1;
  Setup s = make_setup(cfg, 0, 0, 0, 0);
  for (int d = 0; d < 3; ++d) L[d] = s.L[d];
  return 0;
}

// Seeded uniform draws in [lo, hi) for an arbitrary stream (weights,
// finite-difference directions ...).
void sg_uniform(uint64_t seed, uint64_t stream, double lo, double hi, int64_t n, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < n; ++q) out[q] = lo + (hi - lo) * u01(seed, stream, (uint64_t)q);
}

int sg_generate(int cfg, int I_, int J_, int K_, uint64_t seed, double dt, sg_fields* F) {
  if (cfg < 0 || cfg > 4 || !F) return -1;
  Setup s = make_setup(cfg, I_, J_, K_, seed);
  const int I = s.I, J = s.J, K = s.K;
  if (I < 2 || J < 2 || K < 2) return -2;
  Grid1 gx, gy, gz;
  build_uniform(gx, I, s.L[0]);
  build_uniform(gy, J, s.L[1]);
  if (s.zstretch) build_stretched(gz, K, s.L[2], 0.6, 1.02);
  else build_uniform(gz, K, s.L[2]);
  finish_grid(gx); finish_grid(gy); finish_grid(gz);
  for (int i = 1; i <= I; ++i) { F->dxp[i - 1] = gx.d[i]; F->dxu[i - 1] = gx.dc[i]; F->xc[i - 1] = gx.c[i]; }
  for (int j = 1; j <= J; ++j) { F->dyp[j - 1] = gy.d[j]; F->dyv[j - 1] = gy.dc[j]; F->yc[j - 1] = gy.c[j]; }
  for (int k = 1; k <= K; ++k) { F->dzp[k - 1] = gz.d[k]; F->dzw[k - 1] = gz.dc[k]; F->zc[k - 1] = gz.c[k]; }

  const int64_t sx = 1, sy = I + 2, sz = (int64_t)(I + 2) * (J + 2);
  const int64_t Nh = sz * (K + 2);
  auto H = [&](int i, int j, int k) { return (int64_t)i * sx + (int64_t)j * sy + (int64_t)k * sz; };
  auto N3 = [&](int i, int j, int k) { return (int64_t)(i - 1) + (int64_t)I * ((j - 1) + (int64_t)J * (k - 1)); };

  double* arrs[] = {F->u_star, F->v_star, F->w_star, F->ee_area, F->ne_area, F->te_area,
                    F->vof, F->vof_old, F->vof_l, F->phi_l};
  for (double* a : arrs) std::memset(a, 0, sizeof(double) * Nh);
  for (int64_t q = 0; q < Nh; ++q) { F->dp_u[q] = dt; F->dp_v[q] = dt; F->dp_w[q] = dt; }

  // 1. P-cell fractions and level set, halo included
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int k = 0; k <= K + 1; ++k)
    for (int j = 0; j <= J + 1; ++j)
      for (int i = 0; i <= I + 1; ++i) {
        double lo[3], hi[3];
        cell_span(gx, i, lo[0], hi[0]);
        cell_span(gy, j, lo[1], hi[1]);
        cell_span(gz, k, lo[2], hi[2]);
        double fl, lq;
        box_fractions(s, lo, hi, fl, lq);
        int64_t h = H(i, j, k);
        F->vof[h] = fl;
        F->vof_l[h] = lq;
        F->phi_l[h] = gz.c[k] - eta(s, gx.c[i], gy.c[j]);
      }

  // 2. face apertures (fluid fraction of the +x/+y/+z face of each cell)
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int k = 0; k <= K; ++k)
    for (int j = 0; j <= J; ++j)
      for (int i = 0; i <= I; ++i) {
        double lo[3], hi[3];
        if (j >= 1 && k >= 1) {
          cell_span(gy, j, lo[1], hi[1]); cell_span(gz, k, lo[2], hi[2]);
          lo[0] = hi[0] = gx.f[i];
          F->ee_area[H(i, j, k)] = face_fraction(s, 0, gx.f[i], lo, hi);
        }
        if (i >= 1 && k >= 1) {
          cell_span(gx, i, lo[0], hi[0]); cell_span(gz, k, lo[2], hi[2]);
          lo[1] = hi[1] = gy.f[j];
          F->ne_area[H(i, j, k)] = face_fraction(s, 1, gy.f[j], lo, hi);
        }
        if (i >= 1 && j >= 1) {
          cell_span(gx, i, lo[0], hi[0]); cell_span(gy, j, lo[1], hi[1]);
          lo[2] = hi[2] = gz.f[k];
          F->te_area[H(i, j, k)] = face_fraction(s, 2, gz.f[k], lo, hi);
        }
      }

  // 3. DefineCell on every P cell (halo values normalised the same way)
  std::vector<int32_t> ctp((size_t)I * J * K, 0);
  for (int64_t h = 0; h < Nh; ++h) {
    double& v = F->vof[h];
    double& vl = F->vof_l[h];
    if (v >= 1.0 - kEpsi) { vl = vl / v; v = 1.0; }
    if (v <= kEpsi) { v = 0.0; vl = 0.0; }
  }
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        double v = F->vof[H(i, j, k)];
        ctp[N3(i, j, k)] = (v >= 1.0 - kEpsi) ? 0 : (v <= kEpsi ? 2 : 1);
      }
  // 4. CellGeoCal rules: tiny-face demotion, zero solid faces, close cut faces next to solids
  for (int k = 2; k <= K; ++k)
    for (int j = 2; j <= J; ++j)
      for (int i = 2; i <= I; ++i) {
        double mf = std::max({F->ee_area[H(i - 1, j, k)], F->ee_area[H(i, j, k)],
                              F->ne_area[H(i, j - 1, k)], F->ne_area[H(i, j, k)],
                              F->te_area[H(i, j, k - 1)], F->te_area[H(i, j, k)]});
        if (mf < 1.0e-2) ctp[N3(i, j, k)] = 2;
      }
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        int t = ctp[N3(i, j, k)];
        if (t == 2) {
          F->ee_area[H(i - 1, j, k)] = 0; F->ee_area[H(i, j, k)] = 0;
          F->ne_area[H(i, j - 1, k)] = 0; F->ne_area[H(i, j, k)] = 0;
          F->te_area[H(i, j, k - 1)] = 0; F->te_area[H(i, j, k)] = 0;
        }
        if (t == 1) {
          if (i + 1 <= I && ctp[N3(i + 1, j, k)] == 2) F->ee_area[H(i, j, k)] = 0;
          if (j + 1 <= J && ctp[N3(i, j + 1, k)] == 2) F->ne_area[H(i, j, k)] = 0;
          if (k + 1 <= K && ctp[N3(i, j, k + 1)] == 2) F->te_area[H(i, j, k)] = 0;
        }
      }

  // 5. staggered U/V/W cells: types from their own fluid fraction, then MoEx
  std::vector<int32_t> ctu((size_t)I * J * K), ctv((size_t)I * J * K), ctw((size_t)I * J * K);
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        double lo[3], hi[3], fl, lq;
        int32_t* outs[3] = {ctu.data(), ctv.data(), ctw.data()};
        for (int d = 0; d < 3; ++d) {
          cell_span(gx, i, lo[0], hi[0]);
          cell_span(gy, j, lo[1], hi[1]);
          cell_span(gz, k, lo[2], hi[2]);
          const Grid1& g = d == 0 ? gx : (d == 1 ? gy : gz);
          int n = d == 0 ? i : (d == 1 ? j : k);
          lo[d] = g.c[n];
          hi[d] = g.c[n] + g.dc[n];
          box_fractions(s, lo, hi, fl, lq);
          outs[d][N3(i, j, k)] = (fl >= 1.0 - kEpsi) ? 0 : (fl <= kEpsi ? 2 : 1);
        }
      }
  std::vector<int32_t> mu((size_t)I * J * K, 0), mv((size_t)I * J * K, 0), mw((size_t)I * J * K, 0);
  for (int i = 1; i <= I; ++i)
    for (int j = 1; j <= J; ++j)
      for (int k = 1; k <= K; ++k) {
        int64_t n = N3(i, j, k);
        bool ps = ctp[n] == 2;
        if (ctu[n] == 1) { if (i < I) { if (ps || ctp[N3(i + 1, j, k)] == 2) mu[n] = 1; } else if (ps) mu[n] = 1; }
        if (ctv[n] == 1) { if (j < J) { if (ps || ctp[N3(i, j + 1, k)] == 2) mv[n] = 1; } else if (ps) mv[n] = 1; }
        if (ctw[n] == 1) { if (k < K) { if (ps || ctp[N3(i, j, k + 1)] == 2) mw[n] = 1; } else if (ps) mw[n] = 1; }
      }
  // 6. close faces whose velocity the correction forces to zero
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        int64_t n = N3(i, j, k);
        if (ctu[n] == 2 || mu[n] == 1) F->ee_area[H(i, j, k)] = 0;
        if (ctv[n] == 2 || mv[n] == 1) F->ne_area[H(i, j, k)] = 0;
        if (ctw[n] == 2 || mw[n] == 1) F->te_area[H(i, j, k)] = 0;
      }
  // 7. seeded intermediate velocity on interior faces
  const double A = s.ustar_amp;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k <= K; ++k)
    for (int j = 1; j <= J; ++j)
      for (int i = 1; i <= I; ++i) {
        int64_t n = N3(i, j, k), h = H(i, j, k);
        double uu = -A + 2 * A * u01(seed, 1, (uint64_t)n);
        double vv = -A + 2 * A * u01(seed, 2, (uint64_t)n);
        double ww = -A + 2 * A * u01(seed, 3, (uint64_t)n);
        F->u_star[h] = (i < I && F->ee_area[h] != 0.0 && ctu[n] != 2) ? uu : 0.0;
        F->v_star[h] = (j < J && F->ne_area[h] != 0.0 && ctv[n] != 2) ? vv : 0.0;
        F->w_star[h] = (k < K && F->te_area[h] != 0.0 && ctw[n] != 2) ? ww : 0.0;
      }
  std::memcpy(F->vof_old, F->vof, sizeof(double) * Nh);
  std::memcpy(F->cell_type_p, ctp.data(), sizeof(int32_t) * ctp.size());
  std::memcpy(F->cell_type_u, ctu.data(), sizeof(int32_t) * ctu.size());
  std::memcpy(F->cell_type_v, ctv.data(), sizeof(int32_t) * ctv.size());
  std::memcpy(F->cell_type_w, ctw.data(), sizeof(int32_t) * ctw.size());
  std::memcpy(F->moex_u, mu.data(), sizeof(int32_t) * mu.size());
  std::memcpy(F->moex_v, mv.data(), sizeof(int32_t) * mv.size());
  std::memcpy(F->moex_w, mw.data(), sizeof(int32_t) * mw.size());
  return 0;
}

}  // extern "C"

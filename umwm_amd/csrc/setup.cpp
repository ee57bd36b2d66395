// Host-side model setup (see setup.hpp).  Each step cites the reference
// routine whose result it reproduces; arithmetic keeps the reference's
// expression order so the tables equal a serial FP64 build of it.
#include "setup.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <thread>

namespace umwm {
namespace {

constexpr double kPi = 3.141592653589793;  // src/umwm_module.F90:87
constexpr double kTwoPi = 2 * kPi;
constexpr double kREarth = 6.371009e6;     // src/umwm_init.F90:392

int round_up(int v, int m) { return (v + m - 1) / m * m; }

template <typename F>
void parallel_for(int n, F&& fn) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  int nt = (int)std::min<unsigned>(hw, 16u);
  if (n < 4096 || nt == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  int chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto& t : th) t.join();
}

// src/umwm_util.F90:211-232
double haversine(double lon1, double lon2, double lat1, double lat2) {
  double dlon = std::fabs(lon2 - lon1), dlat = std::fabs(lat2 - lat1);
  double s1 = std::sin(0.5 * dlat), s2 = std::sin(0.5 * dlon);
  return 2 * std::asin(std::sqrt(s1 * s1 + std::cos(lat1) * std::cos(lat2) * (s2 * s2)));
}

double bearing(double lo1, double la1, double lo2, double la2) {
  return std::atan2(std::sin(lo2 - lo1) * std::cos(la2),
                    std::cos(la1) * std::sin(la2) -
                        std::sin(la1) * std::cos(la2) * std::cos(lo2 - lo1));
}

struct Grid2 {
  int mm;
  double& at(std::vector<double>& a, int m, int n) const {
    return a[(size_t)(n - 1) * mm + (m - 1)];
  }
  int& at(std::vector<int>& a, int m, int n) const {
    return a[(size_t)(n - 1) * mm + (m - 1)];
  }
};

// src/umwm_init.F90:381-568
void grid(HostSetup& S, const umwm_grid_input* g) {
  const umwm_params& P = S.P;
  const int mm = S.mm, nm = S.nm;
  const size_t n2 = (size_t)mm * nm;
  Grid2 G{mm};
  S.lon.assign(n2, 0.0);
  S.lat.assign(n2, 0.0);
  S.dx2.assign(n2, 0.0);
  S.dy2.assign(n2, 0.0);
  S.curv2.assign(n2, 0.0);
  S.d2.assign(n2, 0.0);
  if (P.gridfromfile) {
    std::vector<double> rlon(n2), rlat(n2);
    for (size_t q = 0; q < n2; ++q) {
      S.lon[q] = g->lon[q];
      S.lat[q] = g->lat[q];
      rlon[q] = S.lon[q] * kTwoPi / 360.;
      rlat[q] = S.lat[q] * kTwoPi / 360.;
    }
    for (int n = 1; n <= nm; ++n) {
      for (int m = 2; m <= mm - 1; ++m)
        G.at(S.dx2, m, n) =
            kREarth * haversine(0.5 * (G.at(rlon, m - 1, n) + G.at(rlon, m, n)),
                                0.5 * (G.at(rlon, m, n) + G.at(rlon, m + 1, n)),
                                0.5 * (G.at(rlat, m - 1, n) + G.at(rlat, m, n)),
                                0.5 * (G.at(rlat, m, n) + G.at(rlat, m + 1, n)));
      G.at(S.dx2, 1, n) = 2 * G.at(S.dx2, 2, n) - G.at(S.dx2, 3, n);
      G.at(S.dx2, mm, n) = 2 * G.at(S.dx2, mm - 1, n) - G.at(S.dx2, mm - 2, n);
    }
    for (int n = 2; n <= nm - 1; ++n)
      for (int m = 1; m <= mm; ++m)
        G.at(S.dy2, m, n) =
            kREarth * haversine(0.5 * (G.at(rlon, m, n - 1) + G.at(rlon, m, n)),
                                0.5 * (G.at(rlon, m, n) + G.at(rlon, m, n + 1)),
                                0.5 * (G.at(rlat, m, n - 1) + G.at(rlat, m, n)),
                                0.5 * (G.at(rlat, m, n) + G.at(rlat, m, n + 1)));
    for (int m = 1; m <= mm; ++m) {
      G.at(S.dy2, m, 1) = 2 * G.at(S.dy2, m, 2) - G.at(S.dy2, m, 3);
      G.at(S.dy2, m, nm) = 2 * G.at(S.dy2, m, nm - 1) - G.at(S.dy2, m, nm - 2);
    }
    for (int n = 1; n <= nm; ++n) {
      for (int m = 2; m <= mm - 1; ++m)
        G.at(S.curv2, m, n) = bearing(G.at(rlon, m - 1, n), G.at(rlat, m - 1, n),
                                      G.at(rlon, m + 1, n), G.at(rlat, m + 1, n));
      if (P.isglobal) {
        G.at(S.curv2, 1, n) = bearing(G.at(rlon, mm, n), G.at(rlat, mm, n),
                                      G.at(rlon, 2, n), G.at(rlat, 2, n));
        G.at(S.curv2, mm, n) = bearing(G.at(rlon, mm - 1, n), G.at(rlat, mm - 1, n),
                                       G.at(rlon, 1, n), G.at(rlat, 1, n));
      } else {
        G.at(S.curv2, 1, n) = G.at(S.curv2, 2, n);
        G.at(S.curv2, mm, n) = G.at(S.curv2, mm - 1, n);
      }
      for (int m = 1; m <= mm; ++m) G.at(S.curv2, m, n) = G.at(S.curv2, m, n) - 0.5 * kPi;
    }
  } else {
    for (size_t q = 0; q < n2; ++q) {
      S.dx2[q] = P.delx;
      S.dy2[q] = P.dely;
    }
  }
  for (size_t q = 0; q < n2; ++q) S.d2[q] = P.topofromfile ? g->z[q] : P.dpt;
}

// src/umwm_init.F90:698-718 fill, with an explicit stack.
void fill_lake(HostSetup& S, int m0, int n0) {
  Grid2 G{S.mm};
  std::vector<std::pair<int, int>> st{{m0, n0}};
  while (!st.empty()) {
    auto [a, b] = st.back();
    st.pop_back();
    if (a == 1 || a == S.mm || b == 1 || b == S.nm) continue;
    if (G.at(S.mask, a, b) == 0 && !(a == m0 && b == n0)) continue;
    G.at(S.mask, a, b) = 0;
    const int nb[4][2] = {{a - 1, b}, {a + 1, b}, {a, b - 1}, {a, b + 1}};
    for (int q = 3; q >= 0; --q)
      if (G.at(S.mask, nb[q][0], nb[q][1]) == 1) st.push_back({nb[q][0], nb[q][1]});
  }
}

// src/umwm_init.F90:571-694
void masks(HostSetup& S, const umwm_grid_input* g) {
  const umwm_params& P = S.P;
  const int mm = S.mm, nm = S.nm;
  Grid2 G{mm};
  S.mask.assign((size_t)mm * nm, 1);
  for (int m = 1; m <= mm; ++m) G.at(S.mask, m, 1) = G.at(S.mask, m, nm) = 0;
  if (!P.isglobal)
    for (int n = 1; n <= nm; ++n) G.at(S.mask, 1, n) = G.at(S.mask, mm, n) = 0;
  if (P.topofromfile) {
    for (size_t q = 0; q < S.mask.size(); ++q) {
      if (S.d2[q] >= 0) {
        S.mask[q] = 0;
        S.d2[q] = P.dmin;
      }
      S.d2[q] = std::fabs(S.d2[q]);
      if (S.d2[q] < P.dmin) S.d2[q] = P.dmin;
    }
    if (P.fillestuaries) {
      bool iterate = true;
      while (iterate) {
        iterate = false;
        for (int n = 2; n <= nm - 1; ++n)
          for (int m = 2; m <= mm - 1; ++m) {
            if (G.at(S.mask, m, n) != 1) continue;
            int cnt = (G.at(S.mask, m - 1, n) == 1) + (G.at(S.mask, m + 1, n) == 1) +
                      (G.at(S.mask, m, n - 1) == 1) + (G.at(S.mask, m, n + 1) == 1);
            if (cnt <= 1) {
              G.at(S.mask, m, n) = 0;
              iterate = true;
            }
          }
      }
    }
    if (P.filllakes && g && g->exclude)
      for (int q = 0; q < g->nexclude; ++q) fill_lake(S, g->exclude[2 * q], g->exclude[2 * q + 1]);
  } else {
    for (size_t q = 0; q < S.mask.size(); ++q)
      if (S.mask[q] == 0) S.d2[q] = P.dmin;
  }
  S.im = (int)std::count(S.mask.begin(), S.mask.end(), 1);
  S.imm = mm * nm;
}

// src/umwm_init.F90:875-1237 (serial, remap_dir = 'v')
int remap(HostSetup& S, std::string& err) {
  const int mm = S.mm, nm = S.nm, imm = S.imm, im = S.im;
  Grid2 G{mm};
  S.ii.assign((size_t)imm, 0);
  S.dx.assign(imm, 0.0);
  S.dy.assign(imm, 0.0);
  S.d.assign(imm, 0.0);
  int i = 0;
  for (int pass = 1; pass >= 0; --pass)  // sea points first, then land
    for (int m = 1; m <= mm; ++m)
      for (int n = 1; n <= nm; ++n)
        if (G.at(S.mask, m, n) == pass) {
          G.at(S.ii, m, n) = ++i;
          S.dx[i - 1] = G.at(S.dx2, m, n);
          S.dy[i - 1] = G.at(S.dy2, m, n);
          S.d[i - 1] = G.at(S.d2, m, n);
        }
  S.mi.assign(imm, 0);
  S.ni.assign(imm, 0);
  for (int n = 1; n <= nm; ++n)
    for (int m = 1; m <= mm; ++m) {
      int q = G.at(S.ii, m, n) - 1;
      S.mi[q] = m;
      S.ni[q] = n;
    }
  // neighbours of every interior point, 1-based (0 = unset)
  std::vector<int> ns(imm + 1, 0), nn(imm + 1, 0), nw(imm + 1, 0), ne(imm + 1, 0);
  for (int n = 2; n <= nm - 1; ++n)
    for (int m = 2; m <= mm - 1; ++m) {
      int c = G.at(S.ii, m, n);
      ns[c] = G.at(S.ii, m, n - 1);
      nn[c] = G.at(S.ii, m, n + 1);
      nw[c] = G.at(S.ii, m - 1, n);
      ne[c] = G.at(S.ii, m + 1, n);
    }
  if (S.P.isglobal)
    for (int n = 2; n <= nm - 1; ++n) {
      int c = G.at(S.ii, 1, n);
      ns[c] = G.at(S.ii, 1, n - 1);
      nn[c] = G.at(S.ii, 1, n + 1);
      nw[c] = G.at(S.ii, mm, n);
      ne[c] = G.at(S.ii, 2, n);
      c = G.at(S.ii, mm, n);
      ns[c] = G.at(S.ii, mm, n - 1);
      nn[c] = G.at(S.ii, mm, n + 1);
      nw[c] = G.at(S.ii, mm - 1, n);
      ne[c] = G.at(S.ii, 1, n);
    }
  S.ie.assign(im, 0); S.iw.assign(im, 0); S.in.assign(im, 0); S.is.assign(im, 0);
  S.iie.assign(im, 0); S.iiw.assign(im, 0); S.iin.assign(im, 0); S.iis.assign(im, 0);
  S.dxn.assign(im, 0.0); S.dxs.assign(im, 0.0); S.dye.assign(im, 0.0); S.dyw.assign(im, 0.0);
  for (int j = 0; j < im; ++j) {
    const int r = j + 1;  // reference index
    if (ns[r] < 1 || nn[r] < 1 || nw[r] < 1 || ne[r] < 1) {
      err = "sea point without neighbours (mask ring broken)";
      return UMWM_ERR_ARGUMENT;
    }
    S.iis[j] = ns[r] - 1; S.iin[j] = nn[r] - 1; S.iiw[j] = nw[r] - 1; S.iie[j] = ne[r] - 1;
    // cell edges in x and y (:1016-1022)
    S.dxn[j] = 0.5 * (S.dx[j] + S.dx[S.iin[j]]);
    S.dxs[j] = 0.5 * (S.dx[j] + S.dx[S.iis[j]]);
    S.dye[j] = 0.5 * (S.dy[j] + S.dy[S.iie[j]]);
    S.dyw[j] = 0.5 * (S.dy[j] + S.dy[S.iiw[j]]);
    // land / off-tile neighbours alias the ghost (:1195-1200)
    auto alias = [im](int ref) { return (ref < 1 || ref > im) ? im : ref - 1; };
    S.is[j] = alias(ns[r]); S.in[j] = alias(nn[r]); S.iw[j] = alias(nw[r]); S.ie[j] = alias(ne[r]);
  }
  return UMWM_OK;
}

// src/umwm_init.F90:1422-1658
void dispersion(HostSetup& S, double tol) {
  const umwm_params& P = S.P;
  const int om = S.om, im = S.im, ldi = S.ldi;
  const double g = P.g;
  auto T = [ldi](std::vector<double>& a, int o, int j) -> double& { return a[(size_t)o * ldi + j]; };
  const size_t nt = (size_t)om * ldi;
  for (auto* v : {&S.k, &S.kdk, &S.k3dk, &S.k4, &S.oneoverk4, &S.cp0, &S.cg0, &S.invcp0, &S.fkovg,
                  &S.logl2overz, &S.cothkd, &S.sbf, &S.sdv, &S.snl_arg, &S.bf1r, &S.bf2r, &S.dwn})
    v->assign(nt, 0.0);
  std::vector<double> kd(nt, 0.0);
  parallel_for(im, [&](int a, int b) {
    for (int j = a; j < b; ++j) {
      const double dj = S.d[j];
      const double bb = P.sfct / (P.rhow0 * g * (dj * dj));
      const double cpg = kTwoPi * std::sqrt(dj / g);  // deep-water first guess
      for (int o = 0; o < om; ++o) {
        double fn = cpg * S.f[o];
        double kk = fn * fn;
        int counter = 1;
        double dk = 2. * tol;
        while (std::fabs(dk) > tol) {  // Newton-Raphson, :1459-1471
          double t = std::tanh(kk);
          dk = -(fn * fn - kk * t * (1. + bb * kk * kk)) /
               (3. * bb * kk * kk * t + t + kk * (1. + bb * kk * kk) * (1. - t * t));
          kk = kk - dk;
          if (counter == 1000) break;
          counter = counter + 1;
        }
        kk = std::fabs(kk) / dj;
        T(S.k, o, j) = kk;
        double kdv = kk * dj;
        T(kd, o, j) = kdv > 20. ? 20. : kdv;
      }
    }
  });
  for (int o = 0; o < om; ++o)
    for (int j = 0; j <= im; ++j) T(S.cp0, o, j) = T(S.cg0, o, j) = DBL_MIN;
  parallel_for(im, [&](int a, int b) {
    for (int j = a; j < b; ++j)
      for (int o = 0; o < om; ++o) {
        const double kk = T(S.k, o, j), kdv = T(kd, o, j), fo = S.f[o];
        const double cp = kTwoPi * fo / kk;
        T(S.cp0, o, j) = cp;
        T(S.cg0, o, j) = cp * (0.5 + kk * S.d[j] / std::sinh(2. * kdv) +
                               P.sfct * kk * kk / (P.rhow0 * g + P.sfct * kk * kk));
        const double cg = T(S.cg0, o, j);
        T(S.dwn, o, j) = S.dom[o] / std::fabs(cg);
        const double l2 = 0.5 * std::fabs(cp) / fo;
        T(S.k4, o, j) = std::pow(kk, 4.);
        T(S.oneoverk4, o, j) = 1. / T(S.k4, o, j);
        T(S.kdk, o, j) = kk * T(S.dwn, o, j);
        T(S.k3dk, o, j) = std::pow(kk, 3.) * T(S.dwn, o, j);
        T(S.fkovg, o, j) = fo * kk / g;
        T(S.cothkd, o, j) = std::cosh(0.2 * kdv) / std::sinh(0.2 * kdv);
        T(S.invcp0, o, j) = 1. / cp;
        T(S.sbf, o, j) = P.sbf_fac * kk / (std::sinh(2. * kdv)) +
                         P.sbp_fac * kk / (std::cosh(kdv) * std::cosh(kdv));
        T(S.sdv, o, j) = 4. * P.nu_water * std::pow(kk, 2.);
        T(S.logl2overz, o, j) = l2 > 20. ? std::log(20. / P.z) : std::log(l2 / P.z);
      }
  });
  for (int j = 0; j < im; ++j)
    for (int o = 0; o < om - 2; ++o) {
      T(S.bf1r, o, j) = P.snl_fac * S.bf1 * T(S.kdk, o + 1, j) / T(S.kdk, o, j);
      T(S.bf2r, o, j) = P.snl_fac * S.bf2 * T(S.kdk, o + 2, j) / T(S.kdk, o, j);
      T(S.snl_arg, o, j) = 1. - (T(S.bf1r, o, j) + T(S.bf2r, o, j));
    }
  // ghost column: per-frequency minimum over sea points (:1652-1656)
  for (int o = 0; o < om; ++o) {
    double a = INFINITY, b = INFINITY;
    for (int j = 0; j < im; ++j) {
      a = std::min(a, T(S.cp0, o, j));
      b = std::min(b, T(S.cg0, o, j));
    }
    T(S.cp0, o, im) = a;
    T(S.cg0, o, im) = b;
  }
}

void remap_mn2i(const HostSetup& S, const double* fmn, std::vector<double>& out) {
  out.assign(S.ldq, 0.0);
  for (int n = 1; n <= S.nm; ++n)
    for (int m = 1; m <= S.mm; ++m) {
      size_t q2 = (size_t)(n - 1) * S.mm + (m - 1);
      out[S.ii[q2] - 1] = fmn[q2];
    }
}

}  // namespace

std::string check_params(const umwm_params& P) {
  if (P.mm < 3 || P.nm < 3) return "bad value in main.nml: mm and nm must be > 2";
  if (P.pm % 4 != 0 || P.pm < 4) return "bad value in main.nml: pm must be divisible by 4";
  if (P.om < 3) return "bad value in main.nml: om must be > 2";
  if (P.om > 512) return "bad value in main.nml: om must be <= 512";
  if (P.pm > 512) return "bad value in main.nml: pm must be <= 512";
  if (P.fmin <= 0 || P.fmax <= 0 || P.fprog <= 0)
    return "bad value in main.nml: fmin, fmax and fprog must be > 0";
  if (P.fmin >= P.fmax) return "bad value in main.nml: fmin must be < fmax";
  if (P.dtg <= 0) return "bad value in main.nml: dtg must be > 0";
  if (P.z <= 0) return "bad value in main.nml: z must be > 0";
  if (P.gustiness != 0)
    return "bad value in main.nml: gustiness must be 0 (random gusts are not supported)";
  if (P.dmin <= 0) return "bad value in main.nml: dmin must be > 0";
  if (!P.gridfromfile && (P.delx <= 0 || P.dely <= 0))
    return "bad value in main.nml: delx and dely must be > 0";
  if (!P.topofromfile && P.dpt <= 0) return "bad value in main.nml: dpt must be > 0";
  return "";
}

void load_forcing_level(HostSetup& S, const umwm_forcing_input& f) {
  const size_t n2 = (size_t)S.mm * S.nm;
  const umwm_params& P = S.P;
  if (P.winds) {
    S.uwf.assign(f.uw, f.uw + n2);
    S.vwf.assign(f.vw, f.vw + n2);
  }
  if (P.currents) {
    S.ucf.assign(f.uc, f.uc + n2);
    S.vcf.assign(f.vc, f.vc + n2);
    for (size_t q = 0; q < n2; ++q)
      if (S.mask[q] == 0) S.ucf[q] = S.vcf[q] = 0;  // src/umwm_io.F90:64-67
  }
  if (P.seaice) S.ficef.assign(f.fice, f.fice + n2);
  if (P.air_density) remap_mn2i(S, f.rhoa, S.rhoaf);
  if (P.water_density) remap_mn2i(S, f.rhow, S.rhowf);
}

int build_setup(const umwm_params& P, const umwm_grid_input* g, const umwm_forcing_input* f0,
                HostSetup& S, std::string& err) {
  err = check_params(P);
  if (!err.empty()) return UMWM_ERR_ARGUMENT;
  if ((P.gridfromfile && (!g || !g->lon || !g->lat)) || (P.topofromfile && (!g || !g->z))) {
    err = "grid input missing for gridfromfile/topofromfile";
    return UMWM_ERR_ARGUMENT;
  }
  const bool anyforcing = P.winds || P.currents || P.air_density || P.water_density || P.seaice;
  if (anyforcing) {
    if (!f0) {
      err = "forcing snapshot missing for the switched forcing fields";
      return UMWM_ERR_ARGUMENT;
    }
    if ((P.winds && (!f0->uw || !f0->vw)) || (P.currents && (!f0->uc || !f0->vc)) ||
        (P.air_density && !f0->rhoa) || (P.water_density && !f0->rhow) ||
        (P.seaice && !f0->fice)) {
      err = "forcing snapshot lacks a switched field";
      return UMWM_ERR_ARGUMENT;
    }
  }
  S.P = P;
  S.mm = P.mm; S.nm = P.nm; S.om = P.om; S.pm = P.pm;
  S.lm = 0;
  for (int q = 0; q < 100; ++q) S.lm += P.depths[q] >= 0;
  S.depth_l.resize(S.lm);
  for (int l = 0; l < S.lm; ++l) S.depth_l[l] = -P.depths[l];

  grid(S, g);
  masks(S, g);
  if (S.im < 1) {
    err = "the mask leaves no sea points";
    return UMWM_ERR_ARGUMENT;
  }
  S.ldi = round_up(S.im + 1, 64);
  S.ldq = round_up(S.imm, 64);
  int rc = remap(S, err);
  if (rc) return rc;

  const int om = S.om, pm = S.pm, im = S.im;
  // ---- input_nc(starttime), src/umwm_io.F90:13-93 ----
  S.wspd.assign(S.ldq, 0.0); S.wdir.assign(S.ldq, 0.0);
  S.uc.assign(S.ldq, 0.0); S.vc.assign(S.ldq, 0.0);
  S.rhoa.assign(S.ldq, 0.0); S.rhow.assign(S.ldq, 0.0); S.fice.assign(S.ldq, 0.0);
  const size_t n2 = (size_t)S.mm * S.nm;
  S.uwf.assign(n2, 0.0); S.vwf.assign(n2, 0.0); S.ucf.assign(n2, P.uc0);
  S.vcf.assign(n2, P.vc0); S.ficef.assign(n2, 0.0);
  S.rhoaf.assign(S.ldq, 0.0); S.rhowf.assign(S.ldq, 0.0);
  for (int q = 0; q < S.imm; ++q) {
    if (!P.winds) { S.wspd[q] = P.wspd0; S.wdir[q] = P.wdir0; }
    if (!P.currents) { S.uc[q] = P.uc0; S.vc[q] = P.vc0; }
    if (!P.seaice) S.fice[q] = P.fice0;
    if (!P.air_density) S.rhoa[q] = P.rhoa0;
    if (!P.water_density) S.rhow[q] = P.rhow0;
  }
  for (size_t q = 0; q < n2; ++q)
    if (S.mask[q] == 0) S.ucf[q] = S.vcf[q] = 0;
  {
    std::vector<double> ra(n2, P.rhoa0), rw(n2, P.rhow0);
    remap_mn2i(S, ra.data(), S.rhoaf);
    remap_mn2i(S, rw.data(), S.rhowf);
  }
  if (anyforcing) load_forcing_level(S, *f0);

  // ---- init, src/umwm_init.F90:1240-1419 ----
  S.dlnf = (std::log(P.fmax) - std::log(P.fmin)) / (double)(om - 1);
  S.f.resize(om);
  S.dom.resize(om);
  for (int o = 0; o < om; ++o) S.f[o] = std::exp(std::log(P.fmin) + o * S.dlnf);
  S.dth = kTwoPi / (double)pm;
  S.dthg = S.dth * P.g;
  S.oneovdth = 1. / S.dth;
  S.twopisds_fac = kTwoPi * P.sds_fac;
  S.fieldscale1 = P.sin_diss1 / P.sin_fac;
  S.fieldscale2 = P.sin_diss2 / P.sin_diss1;
  S.inv_sds_power = 1. / P.sds_power;
  S.cfllim = (pm % 8 == 0) ? std::cos(0.25 * kPi - 0.5 * S.dth) : 1. / std::sqrt(2.);
  S.ar.resize(im); S.oneovar.resize(im); S.oneovdx.resize(im); S.oneovdy.resize(im);
  for (int j = 0; j < im; ++j) {
    S.ar[j] = S.dx[j] * S.dy[j];
    S.oneovdx[j] = 1. / S.dx[j];
    S.oneovdy[j] = 1. / S.dy[j];
    S.oneovar[j] = 1. / S.ar[j];
  }
  {
    double b1 = std::exp(-16 * S.dlnf * S.dlnf);
    double b2 = std::exp(-64 * S.dlnf * S.dlnf);
    S.bf1 = b1 / (b1 + b2);
    S.bf2 = b2 / (b1 + b2);
  }
  S.th.resize(pm); S.cth.resize(pm); S.sth.resize(pm); S.c2th.resize(pm); S.s2th.resize(pm);
  S.pl.resize(pm); S.pr.resize(pm);
  for (int p = 0; p < pm; ++p) {
    S.th[p] = ((p + 1) - 0.5 * (pm + 1)) * S.dth;
    S.cth[p] = std::cos(S.th[p]);
    S.sth[p] = std::sin(S.th[p]);
    S.c2th[p] = std::cos(2 * S.th[p]);
    S.s2th[p] = std::sin(2 * S.th[p]);
    S.pl[p] = (p + 1) % pm;
    S.pr[p] = (p + pm - 1) % pm;
  }
  S.cth_curv.assign((size_t)pm * S.ldi, 0.0);
  S.sth_curv.assign((size_t)pm * S.ldi, 0.0);
  S.ccurv.assign(im, 1.0);
  S.scurv.assign(im, 0.0);
  for (int j = 0; j < im; ++j) {
    const double cv = S.curv2[(size_t)(S.ni[j] - 1) * S.mm + (S.mi[j] - 1)];
    S.ccurv[j] = std::cos(cv);
    S.scurv[j] = std::sin(cv);
    for (int p = 0; p < pm; ++p) {
      S.cth_curv[(size_t)p * S.ldi + j] = std::cos(S.th[p] + cv);
      S.sth_curv[(size_t)p * S.ldi + j] = std::sin(S.th[p] + cv);
    }
  }
  for (int o = 0; o < om; ++o) S.dom[o] = kTwoPi * S.dlnf * S.f[o];
  S.cth2pp.assign((size_t)pm * pm, 0.0);
  for (int p = 0; p < pm; ++p)
    for (int pp = 0; pp < pm; ++pp) {
      int ind = (pp - p + pm) % pm;  // cth2(ind) = cos(dth*ind)**2
      double c = std::cos(S.dth * ind);
      S.cth2pp[(size_t)p * pm + pp] = (c * c) * S.dth;
    }
  dispersion(S, 1e-2);
  double mind = INFINITY;
  for (size_t q = 0; q < n2; ++q)
    if (S.mask[q] == 1) mind = std::min(mind, std::min(S.dx2[q], S.dy2[q]));
  S.mindelx = mind;
  double cgmax = -INFINITY;
  for (int o = 0; o < om; ++o)
    for (int j = 0; j < im; ++j) cgmax = std::max(cgmax, S.cg0[(size_t)o * S.ldi + j]);
  S.cgmax = cgmax;
  S.dtamin = 0.98 * S.cfllim * S.mindelx / S.cgmax;
  if (P.winds) {
    std::vector<double> w2(n2);
    for (size_t q = 0; q < n2; ++q) w2[q] = std::sqrt(S.uwf[q] * S.uwf[q] + S.vwf[q] * S.vwf[q]);
    remap_mn2i(S, w2.data(), S.wspd);
  }
  if (P.seaice) remap_mn2i(S, S.ficef.data(), S.fice);
  S.cd.assign(im, 0.0);
  S.ustar.assign(im, 0.0);
  for (int j = 0; j < im; ++j) {
    S.cd[j] = 1.2e-3;
    if (S.wspd[j] > 11) S.cd[j] = (0.49 + 0.065 * S.wspd[j]) * 1e-3;
    S.ustar[j] = std::sqrt(S.cd[j]) * S.wspd[j];
  }
  return UMWM_OK;
}

}  // namespace umwm

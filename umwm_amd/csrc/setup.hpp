// Host-side model setup in double precision: the work umwm_initialize does
// once before the time loop (src/umwm_top.F90:7-29).  Output arrays are in
// the device layout described in common.hpp (0-based, sea points first,
// ghost column at j = im, tables [o][ldi]).
#pragma once

#include <string>
#include <vector>

#include "../../include/umwm_hip.h"

namespace umwm {

struct HostSetup {
  umwm_params P{};
  int mm = 0, nm = 0, om = 0, pm = 0, im = 0, imm = 0, lm = 0;
  int ldi = 0, ldq = 0;

  // 2-D (mm,nm), m fastest
  std::vector<double> lon, lat, dx2, dy2, curv2, d2;
  std::vector<int> mask, ii;  // ii holds the reference's 1-based index

  // unrolled over all imm points (0-based q = reference i - 1)
  std::vector<int> mi, ni;  // 1-based (m,n) of point q
  std::vector<double> dx, dy, d;

  // per sea point j (0..im-1)
  std::vector<int> ie, iw, in, is;      // aliased; ghost = im
  std::vector<int> iie, iiw, iin, iis;  // true; 0-based into imm
  std::vector<double> ar, oneovar, oneovdx, oneovdy, dxn, dxs, dye, dyw;

  // (o, j) tables, [o * ldi + j], ghost column j = im for cp0/cg0
  std::vector<double> k, kdk, k3dk, k4, oneoverk4, cp0, cg0, invcp0, fkovg;
  std::vector<double> logl2overz, cothkd, sbf, sdv, snl_arg, bf1r, bf2r, dwn;
  // (p, j)
  std::vector<double> cth_curv, sth_curv;
  std::vector<double> ccurv, scurv;  // per sea point
  // spectral constants
  std::vector<double> f, th, cth, sth, c2th, s2th, cth2pp, dom, depth_l;
  std::vector<int> pl, pr;  // 0-based

  double dlnf = 0, dth = 0, dthg = 0, oneovdth = 0, twopisds_fac = 0;
  double fieldscale1 = 0, fieldscale2 = 0, inv_sds_power = 0, cfllim = 0;
  double bf1 = 0, bf2 = 0, mindelx = 0, cgmax = 0, dtamin = 0;

  // forcing state after input_nc(start) + init (per unrolled point, ldq)
  std::vector<double> wspd, wdir, uc, vc, rhoa, rhow, fice;
  // forcing level n+1 (2-D or unrolled) when read from snapshots
  std::vector<double> uwf, vwf, ucf, vcf, ficef, rhoaf, rhowf;
  // initial drag coefficient and friction velocity (Large and Pond)
  std::vector<double> cd, ustar;
};

// Validates the namelist values the way nmlread does
// (src/umwm_init.F90:95-187); returns an empty string when valid.
std::string check_params(const umwm_params& P);

// grid + masks + remap + input_nc + init + dispersion.
int build_setup(const umwm_params& P, const umwm_grid_input* g,
                const umwm_forcing_input* f0, HostSetup& S, std::string& err);

// Loads a forcing snapshot into the level n+1 arrays (input_nc for the
// switched fields, src/umwm_io.F90:34-91).
void load_forcing_level(HostSetup& S, const umwm_forcing_input& f);

}  // namespace umwm

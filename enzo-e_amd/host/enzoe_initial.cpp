// enzoe_initial.cpp -- initial values of the reference's VL+CT test problems
// (the inclined linear waves), computed on the host like the reference's
// initializers and copied into a block's device fields.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <vector>

#include "enzoe_host.hpp"

namespace {

const double kPi = 3.14159265358979324;  // cello::pi (src/Cello/cello.hpp:623)

// the rotation taking domain coordinates to the wave frame (x0 along the
// wave): rows of R(alpha, beta) (EnzoInitialInclinedWave.cpp:20-90)
struct Rotation {
  double m[3][3];
  Rotation(double a, double b) {
    m[0][0] = std::cos(a) * std::cos(b);
    m[0][1] = std::cos(a) * std::sin(b);
    m[0][2] = std::sin(a);
    m[1][0] = -1. * std::sin(b);
    m[1][1] = std::cos(b);
    m[1][2] = 0.;
    m[2][0] = -1. * std::sin(a) * std::cos(b);
    m[2][1] = -1. * std::sin(a) * std::sin(b);
    m[2][2] = std::cos(a);
  }
  void rot(double v0, double v1, double v2, double &o0, double &o1, double &o2) const {
    o0 = m[0][0] * v0 + m[0][1] * v1 + m[0][2] * v2;
    o1 = m[1][0] * v0 + m[1][1] * v1 + m[1][2] * v2;
    o2 = m[2][0] * v0 + m[2][1] * v1 + m[2][2] * v2;
  }
  void inv(double v0, double v1, double v2, double &o0, double &o1, double &o2) const {
    o0 = m[0][0] * v0 + m[1][0] * v1 + m[2][0] * v2;
    o1 = m[0][1] * v0 + m[1][1] * v1 + m[2][1] * v2;
    o2 = m[0][2] * v0 + m[1][2] * v1 + m[2][2] * v2;
  }
};

// background + eigenvector of one wave in the wave frame (conserved form)
struct Wave {
  double d_back = 0, d_ev = 0, e_back = 0, e_ev = 0;
  double m_back[3] = {0, 0, 0}, m_ev[3] = {0, 0, 0};
  bool has_a = false;
  double b_back[3] = {0, 0, 0}, b1_ev = 0, b2_ev = 0;
};

bool is_mhd_wave(const std::string &w) {
  return w == "fast" || w == "alfven" || w == "slow" || w == "mhd_entropy";
}

bool is_hd_wave(const std::string &w) {
  return w == "sound" || w == "hd_entropy" || w == "hd_transv_entropy_v1" ||
         w == "hd_transv_entropy_v2";
}

// prepare_MHD_initializers_ (EnzoInitialInclinedWave.cpp:885-1010): the
// Gardiner & Stone (2008) background (B = (1, 1.5, 0) in the wave frame) and
// right eigenvectors
Wave mhd_wave(const std::string &type, double gamma, double wsign) {
  Wave w;
  w.d_back = 1;
  double etot_back = (1. / gamma) / (gamma - 1.) + 1.625;
  if (type == "mhd_entropy") {
    w.m_back[0] = wsign;
    etot_back += 0.5;
  }
  if (type == "fast") {
    const double coef = 0.5 / std::sqrt(5.);
    w.d_ev = 2. * coef;
    w.m_ev[0] = wsign * 4. * coef;
    w.m_ev[1] = -1. * wsign * 2. * coef;
    w.e_ev = 9. * coef;
    w.b1_ev = 4. * coef;
  } else if (type == "alfven") {
    w.m_ev[2] = -1. * wsign * 1;
    w.b2_ev = 1.;
  } else if (type == "slow") {
    const double coef = 0.5 / std::sqrt(5.);
    w.d_ev = 4. * coef;
    w.m_ev[0] = wsign * 2. * coef;
    w.m_ev[1] = wsign * 4. * coef;
    w.e_ev = 3. * coef;
    w.b1_ev = -2. * coef;
  } else {  // mhd_entropy
    const double coef = 0.5;
    w.d_ev = 2. * coef;
    w.m_ev[0] = 2. * coef * wsign;
    w.e_ev = 1. * coef;
  }
  w.e_back = etot_back;
  w.has_a = true;
  w.b_back[0] = 1.;
  w.b_back[1] = 1.5;
  w.b_back[2] = 0.0;
  return w;
}

// prepare_HD_initializers_ (EnzoInitialInclinedWave.cpp:1014-1120)
Wave hd_wave(const std::string &type, double gamma, double wsign) {
  Wave w;
  double v_back[3] = {0, 0, 0};
  if (type != "sound") v_back[0] = wsign;
  const double v2 = v_back[0] * v_back[0] + v_back[1] * v_back[1] + v_back[2] * v_back[2];
  w.d_back = 1;
  if (type == "sound") {
    const double h_back = 1 / (gamma - 1.) + 0.5 * v2;
    const double signed_cs = wsign * 1;
    w.d_ev = 1;
    w.m_ev[0] = v_back[0] + signed_cs;
    w.m_ev[1] = v_back[1];
    w.m_ev[2] = v_back[2];
    w.e_ev = h_back + v_back[0] * signed_cs;
  } else if (type == "hd_entropy") {
    w.d_ev = 1;
    for (int a = 0; a < 3; a++) w.m_ev[a] = v_back[a];
    w.e_ev = 0.5 * v2;
  } else if (type == "hd_transv_entropy_v1") {
    w.m_ev[1] = 1;
    w.e_ev = v_back[1];
  } else {  // hd_transv_entropy_v2
    w.m_ev[2] = 1;
    w.e_ev = v_back[2];
  }
  w.e_back = ((1. / gamma) / (gamma - 1.) + 0.5 * v2);
  for (int a = 0; a < 3; a++) w.m_back[a] = v_back[a];
  return w;
}

void to_device(double *dst, const std::vector<double> &src) {
  if (hipMemcpy(dst, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess)
    ENZOE_ERROR("EnzoInitialInclinedWave::enforce_block", "hipMemcpy failed");
}

}  // namespace

EnzoInitialInclinedWave::EnzoInitialInclinedWave(ParameterGroup p)
    : alpha_(p.value_float("alpha", 0.0)),
      beta_(p.value_float("beta", 0.0)),
      amplitude_(p.value_float("amplitude", 1.e-6)),
      lambda_(p.value_float("lambda", 1.0)),
      pos_vel_(true) {
  const std::string pv = p.value_string("positive_vel", "true");
  ENZOE_ASSERT("EnzoInitialInclinedWave", p.full_name("positive_vel") + " must be true or false",
               pv == "true" || pv == "false");
  pos_vel_ = (pv == "true");
  // :659-666
  if (p.param("wave_type") == nullptr)
    ENZOE_ERROR("EnzoInitialInclinedWave::EnzoInitialInclinedWave",
                p.full_name("wave_type") + " must be specified");
  wave_type_ = p.value_string("wave_type", "");
  ENZOE_ASSERT("EnzoInitialInclinedWave", "parallel_vel is not provided",
               p.param("parallel_vel") == nullptr);
  if (wave_type_ == "circ_alfven" || wave_type_ == "jeans")
    ENZOE_ERROR("EnzoInitialInclinedWave",
                "wave_type \"" + wave_type_ + "\" is not provided by the device library");
  if (!is_mhd_wave(wave_type_) && !is_hd_wave(wave_type_))
    ENZOE_ERROR("EnzoInitialInclinedWave",
                p.full_name("wave_type") + " has invalid value, \"" + wave_type_ +
                    "\". Must be 'fast', 'alfven', 'slow', 'mhd_entropy', "
                    "'circ_alfven', 'sound', 'hd_entropy', 'hd_transv_entropy_v1', "
                    "'hd_transv_entropy_v2', 'jeans'");
}

void EnzoInitialInclinedWave::enforce_block(Block *block,
                                            const std::array<double, 3> &lower) const {
  const bool mhd = block->has_field("bfield_x");
  const double gamma = enzo::fluid_props()->gamma();
  const double wsign = pos_vel_ ? 1. : -1;
  Wave w;
  if (is_hd_wave(wave_type_)) {
    w = hd_wave(wave_type_, gamma, wsign);
  } else {
    ENZOE_ASSERT("EnzoInitialInclinedWave::enforce_block",
                 "A MHD wave requires fields to store magnetic field values.", mhd);
    w = mhd_wave(wave_type_, gamma, wsign);
  }
  const Rotation R(alpha_, beta_);
  const EFlt3DArray rho_v = block->field_view("density");
  const int mz = rho_v.shape(0), my = rho_v.shape(1), mx = rho_v.shape(2);
  const int g = block->ghost_depth();
  const double *cw = block->CellWidth();
  const double dx = cw[0], dy = cw[1], dz = cw[2];
  // MeshPos (EnzoInitialInclinedWave.cpp:455-500): cell centres and faces
  auto PX = [&](int i) { return lower[0] + dx * (0.5 + (double)(i - g)); };
  auto PY = [&](int j) { return lower[1] + dy * (0.5 + (double)(j - g)); };
  auto PZ = [&](int k) { return lower[2] + dz * (0.5 + (double)(k - g)); };
  auto FX = [&](int i) { return lower[0] + dx * (double)(i - g); };
  auto FY = [&](int j) { return lower[1] + dy * (double)(j - g); };
  auto FZ = [&](int k) { return lower[2] + dz * (double)(k - g); };
  auto at = [](int k, int j, int i, int ny, int nx) { return ((long)k * ny + j) * nx + i; };
  // LinearScalarInit with use_cosine (EnzoInitialInclinedWave.cpp:95-140)
  auto scalar = [&](double back, double ev, double x, double y, double z) {
    double x0, x1, x2;
    R.rot(x, y, z, x0, x1, x2);
    const double tmp = x0 * 2. * kPi / lambda_;
    return (back + amplitude_ * ev * (1 * std::cos(tmp) + (1 - 1) * std::sin(tmp)));
  };
  // RotatedVectorInit of a LinearVectorInit (:145-230)
  auto vector = [&](double x, double y, double z, double &v0, double &v1, double &v2) {
    double x0, x1, x2;
    R.rot(x, y, z, x0, x1, x2);
    const double tmp = x0 * 2. * kPi / lambda_;
    const double trig = (1 * std::cos(tmp) + (1 - 1) * std::sin(tmp));
    const double r0 = w.m_back[0] + amplitude_ * w.m_ev[0] * trig;
    const double r1 = w.m_back[1] + amplitude_ * w.m_ev[1] * trig;
    const double r2 = w.m_back[2] + amplitude_ * w.m_ev[2] * trig;
    R.inv(r0, r1, r2, v0, v1, v2);
  };
  // RotatedVectorInit of a LinearVectorPotentialInit (:235-330)
  auto potential = [&](double x, double y, double z, double &a0, double &a1, double &a2) {
    double x0, x1, x2;
    R.rot(x, y, z, x0, x1, x2);
    const double r0 = (x2 * amplitude_ * w.b1_ev * std::cos(2. * kPi * x0 / lambda_) -
                       x1 * amplitude_ * w.b2_ev * std::cos(2. * kPi * x0 / lambda_));
    const double r1 = w.b_back[2] * x0;
    const double r2 = w.b_back[0] * x1 - w.b_back[1] * x0;
    R.inv(r0, r1, r2, a0, a1, a2);
  };

  const std::size_t ncc = (std::size_t)mx * my * mz;
  std::vector<double> cx, cy, cz;  // cell-centred B (MHD)
  if (mhd) {
    std::vector<double> bx((std::size_t)mz * my * (mx + 1), 0.0),
        by((std::size_t)mz * (my + 1) * mx, 0.0), bz((std::size_t)(mz + 1) * my * mx, 0.0);
    if (w.has_a) {
      // setup_bfield (:506-565): the vector potential on cell edges, then
      // EnzoInitialBCenter's face B = curl A (EnzoInitialBCenter.cpp:43-127)
      std::vector<double> Ax((std::size_t)(mz + 1) * (my + 1) * mx, 0.0),
          Ay((std::size_t)(mz + 1) * my * (mx + 1), 0.0),
          Az((std::size_t)mz * (my + 1) * (mx + 1), 0.0);
      for (int k = 0; k < mz + 1; k++)
        for (int j = 0; j < my + 1; j++)
          for (int i = 0; i < mx + 1; i++) {
            double a0, a1, a2;
            if (i != mx) {
              potential(PX(i), FY(j), FZ(k), a0, a1, a2);
              Ax[at(k, j, i, my + 1, mx)] = a0;
            }
            if (j != my) {
              potential(FX(i), PY(j), FZ(k), a0, a1, a2);
              Ay[at(k, j, i, my, mx + 1)] = a1;
            }
            if (k != mz) {
              potential(FX(i), FY(j), PZ(k), a0, a1, a2);
              Az[at(k, j, i, my + 1, mx + 1)] = a2;
            }
          }
      for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
          for (int i = 0; i < mx + 1; i++)
            bx[at(k, j, i, my, mx + 1)] =
                ((Az[at(k, j + 1, i, my + 1, mx + 1)] - Az[at(k, j, i, my + 1, mx + 1)]) / dy -
                 (Ay[at(k + 1, j, i, my, mx + 1)] - Ay[at(k, j, i, my, mx + 1)]) / dz);
      for (int k = 0; k < mz; k++)
        for (int j = 0; j < my + 1; j++)
          for (int i = 0; i < mx; i++)
            by[at(k, j, i, my + 1, mx)] =
                ((Ax[at(k + 1, j, i, my + 1, mx)] - Ax[at(k, j, i, my + 1, mx)]) / dz -
                 (Az[at(k, j, i + 1, my + 1, mx + 1)] - Az[at(k, j, i, my + 1, mx + 1)]) / dx);
      for (int k = 0; k < mz + 1; k++)
        for (int j = 0; j < my; j++)
          for (int i = 0; i < mx; i++)
            bz[at(k, j, i, my, mx)] =
                ((Ay[at(k, j, i + 1, my, mx + 1)] - Ay[at(k, j, i, my, mx + 1)]) / dx -
                 (Ax[at(k, j + 1, i, my + 1, mx)] - Ax[at(k, j, i, my + 1, mx)]) / dy);
    }
    // cell-centred B = face average (EnzoBfieldMethodCT::compute_center_bfield)
    cx.assign(ncc, 0.0);
    cy.assign(ncc, 0.0);
    cz.assign(ncc, 0.0);
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          const long c = at(k, j, i, my, mx);
          cx[c] = 0.5 * (bx[at(k, j, i, my, mx + 1)] + bx[at(k, j, i + 1, my, mx + 1)]);
          cy[c] = 0.5 * (by[at(k, j, i, my + 1, mx)] + by[at(k, j + 1, i, my + 1, mx)]);
          cz[c] = 0.5 * (bz[at(k, j, i, my, mx)] + bz[at(k + 1, j, i, my, mx)]);
        }
    to_device(block->field_view("bfieldi_x").data(), bx);
    to_device(block->field_view("bfieldi_y").data(), by);
    to_device(block->field_view("bfieldi_z").data(), bz);
    to_device(block->field_view("bfield_x").data(), cx);
    to_device(block->field_view("bfield_y").data(), cy);
    to_device(block->field_view("bfield_z").data(), cz);
  }
  // setup_fluid_ in conserved form (:627-660): specific velocity and energy
  std::vector<double> rho(ncc), vx(ncc), vy(ncc), vz(ncc), etot(ncc);
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const double x = PX(i), y = PY(j), z = PZ(k);
        const long c = at(k, j, i, my, mx);
        const double d = scalar(w.d_back, w.d_ev, x, y, z);
        rho[c] = d;
        double px, py, pz;
        vector(x, y, z, px, py, pz);
        vx[c] = px / d;
        vy[c] = py / d;
        vz[c] = pz / d;
        etot[c] = scalar(w.e_back, w.e_ev, x, y, z) / d;
      }
  to_device(block->field_view("density").data(), rho);
  to_device(block->field_view("velocity_x").data(), vx);
  to_device(block->field_view("velocity_y").data(), vy);
  to_device(block->field_view("velocity_z").data(), vz);
  to_device(block->field_view("total_energy").data(), etot);
  // setup_eint_ (:495-538): with an internal_energy field (dual energy),
  // eint = etot - kinetic - magnetic / density
  if (block->has_field("internal_energy")) {
    std::vector<double> eint(ncc);
    for (std::size_t c = 0; c < ncc; c++) {
      const double kinetic = 0.5 * (vx[c] * vx[c] + vy[c] * vy[c] + vz[c] * vz[c]);
      const double magnetic =
          mhd ? 0.5 * (cx[c] * cx[c] + cy[c] * cy[c] + cz[c] * cz[c]) / rho[c] : 0.;
      eint[c] = etot[c] - kinetic - magnetic;
    }
    to_device(block->field_view("internal_energy").data(), eint);
  }
}

// ------------------------------------------------------------ shock tubes

namespace {

// left / right states (EnzoInitialShockTube.cpp:36-86): density, pressure,
// velocity_0..2 (0 along the tube), bfield_1, bfield_2 (transverse)
struct TubeState {
  double density, pressure, v[3], b1, b2;
};

bool tube_states(const std::string &name, TubeState &l, TubeState &r, double &b0) {
  if (name == "rj2a") {
    l = {1.08, 0.95, {1.2, 0.01, 0.5}, 1.0155412503859613, 0.5641895835477563};
    r = {1., 1.0, {0.0, 0.0, 0.0}, 1.1283791670955126, 0.5641895835477563};
    b0 = 0.5641895835477563;
    return true;
  }
  if (name == "sod") {
    l = {1.0, 1.0, {0., 0., 0.}, 0., 0.};
    r = {0.125, 0.1, {0., 0., 0.}, 0., 0.};
    b0 = 0.0;
    return true;
  }
  return false;
}

// setup_maps_ (:161-176): vectors change sign when the tube is flipped
TubeState flipped(const TubeState &s) {
  TubeState o = s;
  for (int a = 0; a < 3; a++) o.v[a] = -1. * s.v[a];
  o.b1 = -1. * s.b1;
  o.b2 = -1. * s.b2;
  return o;
}

}  // namespace

EnzoInitialShockTube::EnzoInitialShockTube(ParameterGroup p)
    : setup_name_(p.value_string("setup_name", "")),
      axis_velocity_(p.value_float("axis_velocity", 0.0)),
      trans_velocity_(p.value_float("transverse_velocity", 0.0)) {
  TubeState l, r;
  double b0 = 0.0;
  // :107-122
  if (!tube_states(setup_name_, l, r, b0))
    ENZOE_ERROR("EnzoInitialShockTube", p.full_name("setup_name") +
                                            " specifies an invalid name: (must be rj2a, sod), "
                                            "not " + setup_name_ + ".");
  const std::string ax = p.value_string("aligned_ax", "x");
  if (ax == "x")
    aligned_ax_ = 0;
  else if (ax == "y")
    aligned_ax_ = 1;
  else if (ax == "z")
    aligned_ax_ = 2;
  else
    ENZOE_ERROR("EnzoInitialShockTube", p.full_name("aligned_ax") +
                                            " must specify \"x\", \"y\", or \"z\", not \"" +
                                            ax + "\"");
  const std::string flip = p.value_string("flip_initialize", "false");
  ENZOE_ASSERT("EnzoInitialShockTube", p.full_name("flip_initialize") + " must be true or false",
               flip == "true" || flip == "false");
  flipped_initialize_ = (flip == "true");
}

void EnzoInitialShockTube::enforce_block(Block *block, const std::array<double, 3> &lower,
                                         const std::array<double, 3> &upper) const {
  TubeState l0, r0;
  double b0 = 0.0;
  tube_states(setup_name_, l0, r0, b0);
  const TubeState L = flipped_initialize_ ? flipped(r0) : l0;
  const TubeState R = flipped_initialize_ ? flipped(l0) : r0;
  const double gamma = enzo::fluid_props()->gamma();
  const double flip = flipped_initialize_ ? -1. : 1.;
  const double aligned_bfield_val = flip * b0;
  const double axis_velocity = flip * axis_velocity_;
  const double trans_velocity = flip * trans_velocity_;
  const bool mhd = block->has_field("bfieldi_x");
  ENZOE_ASSERT("EnzoInitialShockTube::enforce_block",
               "the \"" + setup_name_ + "\" tube needs the magnetic fields",
               mhd || (b0 == 0.0 && L.b1 == 0.0 && L.b2 == 0.0 && R.b1 == 0.0 &&
                       R.b2 == 0.0));

  const EFlt3DArray rho_v = block->field_view("density");
  const int m[3] = {rho_v.shape(2), rho_v.shape(1), rho_v.shape(0)};
  const int g = block->ghost_depth();
  const int ai = aligned_ax_, aj = (ai + 1) % 3, ak = (ai + 2) % 3;
  // prep_aligned_slices_ (:296-335): cells [0, shock_ind) left, the rest right
  const int n_ai = m[ai] - 2 * g;
  const double di = (upper[ai] - lower[ai]) / n_ai;
  int shock_ind = (int)std::ceil((0.5 - lower[ai]) / di - 0.5 + (double)g);
  shock_ind = std::min(std::max(shock_ind, 0), m[ai]);

  const char *vel[3] = {"velocity_x", "velocity_y", "velocity_z"};
  const char *bfi[3] = {"bfieldi_x", "bfieldi_y", "bfieldi_z"};
  // host copies of the fields this initializer writes
  auto download = [&](const char *name) {
    const EFlt3DArray a = block->field_view(name);
    std::vector<double> h((std::size_t)a.shape(0) * a.shape(1) * a.shape(2));
    if (hipMemcpy(h.data(), a.data(), h.size() * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess)
      ENZOE_ERROR("EnzoInitialShockTube::enforce_block", "hipMemcpy failed");
    return h;
  };
  // initializer_helper_ (:338-350): the slice along the tube, every index
  // across it (face-centred extents included)
  auto fill = [&](const char *name, int s0, int s1, double val, std::vector<double> &h) {
    const EFlt3DArray a = block->field_view(name);
    const int ext[3] = {a.shape(2), a.shape(1), a.shape(0)};
    for (int k = 0; k < ext[2]; k++)
      for (int j = 0; j < ext[1]; j++)
        for (int i = 0; i < ext[0]; i++) {
          const int p[3] = {i, j, k};
          if (p[ai] < s0 || p[ai] >= s1) continue;
          h[((std::size_t)k * ext[1] + j) * ext[0] + i] = val;
        }
  };
  std::vector<double> rho = download("density"), etot = download("total_energy");
  std::vector<double> v[3] = {download(vel[0]), download(vel[1]), download(vel[2])};
  const bool has_eint = block->has_field("internal_energy");
  std::vector<double> eint_h;
  if (has_eint) eint_h = download("internal_energy");
  std::vector<double> bf[3];
  if (mhd)
    for (int a = 0; a < 3; a++) bf[a] = download(bfi[a]);
  for (int side = 0; side < 2; side++) {
    const TubeState &S = side ? R : L;
    const int s0 = side ? shock_ind : 0, s1 = side ? m[ai] : shock_ind;
    if (s0 >= s1) continue;
    const double velocity_0 = S.v[0] + axis_velocity;
    const double velocity_1 = S.v[1] + trans_velocity;
    const double velocity_2 = S.v[2];
    fill("density", s0, s1, S.density, rho);
    fill(vel[ai], s0, s1, velocity_0, v[ai]);
    fill(vel[aj], s0, s1, velocity_1, v[aj]);
    fill(vel[ak], s0, s1, velocity_2, v[ak]);
    if (mhd) {
      fill(bfi[aj], s0, s1, S.b1, bf[aj]);
      fill(bfi[ak], s0, s1, S.b2, bf[ak]);
    }
    const double eint = (S.pressure / ((gamma - 1.) * S.density));
    // (optionally) the specific internal energy (:255-261)
    if (has_eint) fill("internal_energy", s0, s1, eint, eint_h);
    const double v2 = (velocity_0 * velocity_0 + velocity_1 * velocity_1 +
                       velocity_2 * velocity_2);
    const double b2 = (aligned_bfield_val * aligned_bfield_val + S.b1 * S.b1 + S.b2 * S.b2);
    fill("total_energy", s0, s1, (eint + 0.5 * (v2 + b2 / S.density)), etot);
  }
  to_device(block->field_view("density").data(), rho);
  to_device(block->field_view("total_energy").data(), etot);
  if (has_eint) to_device(block->field_view("internal_energy").data(), eint_h);
  for (int a = 0; a < 3; a++) to_device(block->field_view(vel[a]).data(), v[a]);
  if (mhd) {
    // the aligned face field is uniform (:279-280)
    std::fill(bf[ai].begin(), bf[ai].end(), aligned_bfield_val);
    for (int a = 0; a < 3; a++) to_device(block->field_view(bfi[a]).data(), bf[a]);
    // EnzoInitialBCenter::initialize_bfield_center: cell-centred B = face average
    const int mx = m[0], my = m[1], mz = m[2];
    std::vector<double> c[3];
    for (int a = 0; a < 3; a++) c[a].assign((std::size_t)mx * my * mz, 0.0);
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          const std::size_t q = ((std::size_t)k * my + j) * mx + i;
          c[0][q] = 0.5 * (bf[0][((std::size_t)k * my + j) * (mx + 1) + i] +
                           bf[0][((std::size_t)k * my + j) * (mx + 1) + i + 1]);
          c[1][q] = 0.5 * (bf[1][((std::size_t)k * (my + 1) + j) * mx + i] +
                           bf[1][((std::size_t)k * (my + 1) + j + 1) * mx + i]);
          c[2][q] = 0.5 * (bf[2][((std::size_t)k * my + j) * mx + i] +
                           bf[2][((std::size_t)(k + 1) * my + j) * mx + i]);
        }
    const char *bc[3] = {"bfield_x", "bfield_y", "bfield_z"};
    for (int a = 0; a < 3; a++) to_device(block->field_view(bc[a]).data(), c[a]);
  }
}

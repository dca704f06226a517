// enzoe_host_capi.cpp -- C entry points over the C++ drop-in classes.
#include "../../include/enzoe_vlct_host.h"

#include <array>
#include <memory>
#include <vector>

#include "enzoe_host.hpp"

struct enzoe_host_method_s {
  std::unique_ptr<EnzoMethodMHDVlct> method;
  std::string name;
};

// a static mesh of hydro blocks under the Method list
// ["mhd_vlct", "flux_correct"]
struct enzoe_host_amr_s {
  std::unique_ptr<EnzoMethodMHDVlct> vlct;
  std::unique_ptr<MethodFluxCorrect> flux_correct;
  std::vector<std::unique_ptr<Block>> blocks;
  struct Link {
    int block, axis, face, coarse;
    std::array<int, 3> child;
  };
  std::vector<Link> links;  // faces whose neighbour is coarser
};

extern "C" {

void enzoe_host_set_fluid_props(double gamma, double density_floor,
                                double pressure_floor) {
  enzo::set_fluid_props(EnzoPhysicsFluidProps(gamma, density_floor, pressure_floor));
}

void enzoe_host_set_dual_energy(const char *type, const double *eta, int n_eta) {
  ENZOE_ASSERT("enzoe_host_set_dual_energy", "n_eta must be 0, 1 or 2",
               n_eta >= 0 && n_eta <= 2 && (n_eta == 0 || eta != nullptr));
  std::vector<double> e;
  for (int i = 0; i < n_eta; i++) e.push_back(eta[i]);
  enzo::fluid_props()->set_dual_energy(type ? type : "disabled", e);
}

enzoe_host_method_t enzoe_host_method_create(const char *parameters) {
  ParameterGroup p(parameters ? parameters : "");
  enzoe_host_method_s *m = new enzoe_host_method_s;
  m->method.reset(new EnzoMethodMHDVlct(p, false));
  m->name = m->method->name();
  return m;
}

void enzoe_host_method_destroy(enzoe_host_method_t m) { delete m; }

const char *enzoe_host_method_name(enzoe_host_method_t m) {
  return m ? m->name.c_str() : "";
}

double enzoe_host_method_courant(enzoe_host_method_t m) {
  return m ? m->method->courant() : -1.0;
}

int enzoe_host_method_compute(enzoe_host_method_t m, const enzoe_block *blocks,
                              int n, double dt, enzoe_stream_t stream) {
  if (!m || !blocks || n <= 0) return 0;
  enzoe_host_set_stream(stream);
  std::vector<std::unique_ptr<Block>> owned;
  std::vector<Block *> list;
  for (int i = 0; i < n; i++) {
    owned.emplace_back(new Block(blocks[i]));
    owned.back()->set_dt(dt);
    list.push_back(owned.back().get());
  }
  if (n == 1)
    m->method->compute(list[0]);
  else
    m->method->compute_blocks(list);
  int done = 0;
  for (Block *b : list) done += b->compute_done_count();
  return done;
}

int enzoe_host_method_compute_passive(enzoe_host_method_t m,
                                      const enzoe_block *block, int npassive,
                                      const char *const *names,
                                      double *const *passive, double dt,
                                      enzoe_stream_t stream) {
  if (!m || !block || npassive < 0 || (npassive > 0 && (!names || !passive))) return 0;
  enzoe_host_set_stream(stream);
  Block b(*block);
  for (int i = 0; i < npassive; i++) b.add_passive_scalar(names[i], passive[i]);
  b.set_dt(dt);
  m->method->compute(&b);
  return b.compute_done_count();
}

int enzoe_host_method_compute_accel(enzoe_host_method_t m, const enzoe_block *block,
                                    double *const *accel, double dt,
                                    enzoe_stream_t stream) {
  if (!m || !block || !accel || !accel[0] || !accel[1] || !accel[2]) return 0;
  enzoe_host_set_stream(stream);
  Block b(*block);
  b.add_field("acceleration_x", accel[0]);
  b.add_field("acceleration_y", accel[1]);
  b.add_field("acceleration_z", accel[2]);
  b.set_dt(dt);
  m->method->compute(&b);
  return b.compute_done_count();
}

double enzoe_host_method_timestep(enzoe_host_method_t m, const enzoe_block *block,
                                  enzoe_stream_t stream) {
  if (!m || !block) return -1.0;
  enzoe_host_set_stream(stream);
  Block b(*block);
  return m->method->timestep(&b);
}

enzoe_host_amr_t enzoe_host_amr_create(const char *vlct_parameters,
                                       const char *flux_correct_parameters,
                                       const enzoe_block *blocks, const int *level,
                                       int n) {
  if (!blocks || !level || n <= 0) return nullptr;
  enzoe_host_amr_s *a = new enzoe_host_amr_s;
  // EnzoProblem.cpp:565-566, 655: flux_correct in the list makes the VL+CT
  // method store its fluxes
  a->vlct.reset(new EnzoMethodMHDVlct(ParameterGroup(vlct_parameters ? vlct_parameters : ""),
                                      true));
  a->flux_correct.reset(new MethodFluxCorrect(
      ParameterGroup(flux_correct_parameters ? flux_correct_parameters : "",
                     "Method:flux_correct")));
  for (int i = 0; i < n; i++) {
    a->blocks.emplace_back(new Block(blocks[i]));
    a->blocks.back()->set_level(level[i]);
  }
  return a;
}

int enzoe_host_amr_set_face(enzoe_host_amr_t a, int block, int axis, int face,
                            int neighbor_level, int coarse_neighbor, const int child[3]) {
  if (!a || block < 0 || block >= (int)a->blocks.size() || axis < 0 || axis > 2 ||
      (face != 0 && face != 1))
    return ENZOE_ERR_ARG;
  a->blocks[block]->set_face_level(axis, face, neighbor_level);
  if (coarse_neighbor >= 0) {
    if (coarse_neighbor >= (int)a->blocks.size() || !child) return ENZOE_ERR_ARG;
    a->links.push_back({block, axis, face, coarse_neighbor, {child[0], child[1], child[2]}});
  }
  return ENZOE_OK;
}

int enzoe_host_amr_cycle(enzoe_host_amr_t a, double dt, enzoe_stream_t stream,
                         double *digits, int ndigits) {
  if (!a) return ENZOE_ERR_ARG;
  enzoe_host_set_stream(stream);
  std::vector<Block *> list;
  for (auto &b : a->blocks) {
    b->set_dt(dt);
    list.push_back(b.get());
  }
  // Method 1: mhd_vlct on every block (one batched launch)
  a->vlct->compute_blocks(list);
  // refresh of the fluxes toward coarser neighbours
  for (const auto &l : a->links)
    refresh_flux_face_to_coarse(a->blocks[l.block].get(), a->blocks[l.coarse].get(), l.axis,
                                l.face, l.child);
  // Method 2: flux_correct over every block and its sum reduction
  a->flux_correct->compute_blocks(list);
  const str_vec_t fields = cello::field_groups()->group_list("conserved");
  for (int i = 0; i < ndigits && i < (int)fields.size(); i++)
    digits[i] = a->flux_correct->conserved_digits(fields[i]);
  return a->flux_correct->min_digits_ok() ? ENZOE_OK : 1;
}

void enzoe_host_amr_destroy(enzoe_host_amr_t a) { delete a; }

int enzoe_host_initial_inclined_wave(const char *parameters, const enzoe_block *block,
                                     const double lower[3]) {
  if (!block || !lower) return ENZOE_ERR_ARG;
  EnzoInitialInclinedWave init(
      ParameterGroup(parameters ? parameters : "", "Initial:inclined_wave"));
  Block b(*block);
  init.enforce_block(&b, {lower[0], lower[1], lower[2]});
  return ENZOE_OK;
}

int enzoe_host_initial_shock_tube(const char *parameters, const enzoe_block *block,
                                  const double lower[3], const double upper[3]) {
  if (!block || !lower || !upper) return ENZOE_ERR_ARG;
  EnzoInitialShockTube init(
      ParameterGroup(parameters ? parameters : "", "Initial:shock_tube"));
  Block b(*block);
  init.enforce_block(&b, {lower[0], lower[1], lower[2]}, {upper[0], upper[1], upper[2]});
  return ENZOE_OK;
}

int enzoe_host_prolong_apply(const char *name, int use_linear, double *values_f,
                             const int m3_f[3], const int o3_f[3], const int n3_f[3],
                             const double *values_c, const int m3_c[3], const int o3_c[3],
                             const int n3_c[3], int accumulate, enzoe_stream_t stream) {
  if (!name) return ENZOE_ERR_ARG;
  enzoe_host_set_stream(stream);
  std::unique_ptr<Prolong> p;
  const std::string n(name);
  if (n == "linear")
    p.reset(new ProlongLinear());
  else if (n == "enzo")
    p.reset(new EnzoProlong("3A", true, use_linear != 0));
  else
    return ENZOE_ERR_CONFIG;
  int mf[3], of[3], nf[3], mc[3], oc[3], nc[3];
  for (int a = 0; a < 3; a++) {
    mf[a] = m3_f[a]; of[a] = o3_f[a]; nf[a] = n3_f[a];
    mc[a] = m3_c[a]; oc[a] = o3_c[a]; nc[a] = n3_c[a];
  }
  p->apply(precision_double, values_f, mf, of, nf, values_c, mc, oc, nc, accumulate != 0);
  return ENZOE_OK;
}

int enzoe_host_enzo_prolong_apply(const char *type, int positive, int use_linear,
                                  double *values_f, const int m3_f[3], const int o3_f[3],
                                  const int n3_f[3], const double *values_c,
                                  const int m3_c[3], const int o3_c[3], const int n3_c[3],
                                  int accumulate, enzoe_stream_t stream) {
  if (!type) return ENZOE_ERR_ARG;
  enzoe_host_set_stream(stream);
  EnzoProlong p(type, positive != 0, use_linear != 0);
  int mf[3], of[3], nf[3], mc[3], oc[3], nc[3];
  for (int a = 0; a < 3; a++) {
    mf[a] = m3_f[a]; of[a] = o3_f[a]; nf[a] = n3_f[a];
    mc[a] = m3_c[a]; oc[a] = o3_c[a]; nc[a] = n3_c[a];
  }
  p.apply(precision_double, values_f, mf, of, nf, values_c, mc, oc, nc, accumulate != 0);
  return ENZOE_OK;
}

int enzoe_host_restrict_apply(const char *name, double *values_c, const int nd3_c[3],
                              const int im3_c[3], const int n3_c[3], const double *values_f,
                              const int nd3_f[3], const int im3_f[3], const int n3_f[3],
                              int accumulate, enzoe_stream_t stream) {
  if (!name) return -1;
  enzoe_host_set_stream(stream);
  std::unique_ptr<Restrict> r;
  const std::string n(name);
  if (n == "linear")
    r.reset(new RestrictLinear());
  else if (n == "enzo")
    r.reset(new EnzoRestrict());
  else
    return -1;
  int ndc[3], imc[3], nc[3], ndf[3], imf[3], nf[3];
  for (int a = 0; a < 3; a++) {
    ndc[a] = nd3_c[a]; imc[a] = im3_c[a]; nc[a] = n3_c[a];
    ndf[a] = nd3_f[a]; imf[a] = im3_f[a]; nf[a] = n3_f[a];
  }
  return r->apply(precision_double, values_c, ndc, imc, nc, values_f, ndf, imf, nf,
                  accumulate != 0);
}

}  // extern "C"

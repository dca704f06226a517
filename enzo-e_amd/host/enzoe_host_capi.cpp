// enzoe_host_capi.cpp -- C entry points over the C++ drop-in classes.
#include "../../include/enzoe_vlct_host.h"

#include <memory>
#include <vector>

#include "enzoe_host.hpp"

struct enzoe_host_method_s {
  std::unique_ptr<EnzoMethodMHDVlct> method;
  std::string name;
};

extern "C" {

void enzoe_host_set_fluid_props(double gamma, double density_floor,
                                double pressure_floor) {
  enzo::set_fluid_props(EnzoPhysicsFluidProps(gamma, density_floor, pressure_floor));
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

double enzoe_host_method_timestep(enzoe_host_method_t m, const enzoe_block *block,
                                  enzoe_stream_t stream) {
  if (!m || !block) return -1.0;
  enzoe_host_set_stream(stream);
  Block b(*block);
  return m->method->timestep(&b);
}

}  // extern "C"

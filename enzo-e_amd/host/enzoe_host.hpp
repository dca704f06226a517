// enzoe_host.hpp -- C++ drop-in layer with the reference's plugin surface for
// the VL+CT method: the same class names, static factories, accepted names,
// argument meaning, defaults and error behaviour as
//   src/Enzo/hydro-mhd/EnzoMethodMHDVlct.{hpp,cpp}
//   src/Enzo/hydro-mhd/EnzoMHDIntegratorStageCommands.{hpp,cpp}
//   src/Enzo/hydro-mhd/toolkit/EnzoReconstructor*.{hpp,cpp}
//   src/Enzo/hydro-mhd/riemann/EnzoRiemann*.{hpp,cpp}
//   src/Enzo/hydro-mhd/toolkit/EnzoBfieldMethod{,CT}.{hpp,cpp}
//   src/Enzo/hydro-mhd/toolkit/EnzoIntegrationQuanUpdate.{hpp,cpp}
// Every array is device memory; every operation is one call of the gfx950
// library through its C ABI (include/enzoe_vlct.h).  The Cello pieces the
// method touches (CelloView, ViewMap, Block, ParameterGroup, the fluid
// properties) are reduced to what the method uses.
#pragma once

#include <array>
#include <cstddef>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/enzoe_vlct.h"

using enzo_float = double;
using str_vec_t = std::vector<std::string>;

// ERROR / ASSERT of the reference (src/Cello/error_Error.hpp:52,264,
// error_Error.cpp:86): print the message and exit(1).
[[noreturn]] void enzoe_host_error(const char *function, const std::string &message);
#define ENZOE_ERROR(F, M) enzoe_host_error((F), (M))
#define ENZOE_ASSERT(F, M, COND) \
  do {                           \
    if (!(COND)) enzoe_host_error((F), (M)); \
  } while (0)

// Stream every plugin call of the calling thread launches on (default: the
// null stream).  One stream per host thread, as one PE per Method instance.
void enzoe_host_set_stream(enzoe_stream_t stream);
enzoe_stream_t enzoe_host_stream();

// ------------------------------------------------------------------ views

// Slice [start, stop) of one axis; a negative bound counts from the end and
// an absent bound is open (src/Cello/view_CSlice.hpp).
class CSlice {
 public:
  CSlice() = default;
  CSlice(std::nullptr_t, std::nullptr_t) {}
  CSlice(int start, int stop) : has_start_(true), has_stop_(true), start_(start), stop_(stop) {}
  CSlice(int start, std::nullptr_t) : has_start_(true), start_(start) {}
  CSlice(std::nullptr_t, int stop) : has_stop_(true), stop_(stop) {}
  // resolved [lo, hi) on an axis of extent n (ERROR when empty or outside)
  void resolve(int n, int *lo, int *hi) const;

 private:
  bool has_start_ = false, has_stop_ = false;
  int start_ = 0, stop_ = 0;
};

// CelloView<double, 3> over device memory (src/Cello/view_CelloView.hpp:
// 356-931): shared ownership of the allocation, offset, (z,y,x) shape and
// strides, row-major with x fastest.
class EFlt3DArray {
 public:
  EFlt3DArray() = default;
  // a new zero-filled device array, owned (shared) by the views
  EFlt3DArray(int mz, int my, int mx);
  // a view of device memory owned elsewhere
  EFlt3DArray(double *data, int mz, int my, int mx);
  EFlt3DArray(std::shared_ptr<void> owner, double *data,
              const std::array<int, 3> &shape,
              const std::array<long long, 3> &stride);

  int shape(unsigned axis) const { return shape_[axis]; }
  long long stride(unsigned axis) const { return stride_[axis]; }
  double *data() const { return data_; }
  bool is_null() const { return data_ == nullptr; }
  std::size_t size() const {
    return (std::size_t)shape_[0] * shape_[1] * shape_[2];
  }
  EFlt3DArray subarray(const CSlice &z, const CSlice &y, const CSlice &x) const;
  enzoe_farray desc() const;

 private:
  std::shared_ptr<void> owner_;
  double *data_ = nullptr;
  std::array<int, 3> shape_{{0, 0, 0}};
  std::array<long long, 3> stride_{{0, 0, 0}};
};

// EnzoPermutedCoordinates (EnzoPermutedCoordinates.hpp:16-104): i = dim,
// j = (dim+1)%3, k = (dim+2)%3; slices given in (k, j, i) order.
class EnzoPermutedCoordinates {
 public:
  explicit EnzoPermutedCoordinates(int i_axis);
  int i_axis() const { return i_; }
  int j_axis() const { return (i_ + 1) % 3; }
  int k_axis() const { return (i_ + 2) % 3; }
  EFlt3DArray get_subarray(const EFlt3DArray &a, const CSlice &k_slice,
                           const CSlice &j_slice, const CSlice &i_slice) const;
  EFlt3DArray left_edge_offset(const EFlt3DArray &a, int k_start, int j_start,
                               int i_start) const;

 private:
  int i_;
};

// EnzoEFltArrayMap = ViewMap<enzo_float> (view_ViewMap.hpp:43-395): keys ->
// same-shape views.  Built from (keys, shape) the views are slices of one
// contiguous (nkeys, mz, my, mx) device array, as ViewCollec
// (view_ViewCollec.hpp:126-170).
class EnzoEFltArrayMap {
 public:
  EnzoEFltArrayMap() = default;
  EnzoEFltArrayMap(std::string name, const str_vec_t &keys,
                   const std::array<int, 3> &shape);
  EnzoEFltArrayMap(std::string name, const str_vec_t &keys,
                   const std::vector<EFlt3DArray> &arrays);

  const std::string &name() const { return name_; }
  std::size_t size() const { return keys_.size(); }
  const str_vec_t &keys() const { return keys_; }
  bool contains(const std::string &key) const;
  EFlt3DArray at(const std::string &key) const;
  EFlt3DArray operator[](const std::string &key) const { return at(key); }
  // view trimmed by stale_depth on every side (view_ViewMap.hpp:310-327)
  EFlt3DArray get(const std::string &key, int stale_depth = 0) const;
  int array_shape(unsigned axis) const;
  EnzoEFltArrayMap subarray_map(const CSlice &z, const CSlice &y,
                                const CSlice &x,
                                const std::string &name = "") const;
  // the device descriptor: views by slot (ENZOE_SLOT_*; the keys of
  // passive_list take slots ENZOE_SLOT_PASSIVE0 + their index); ERROR for a
  // key without a slot or views of different shapes or strides
  enzoe_fmap slot_map(const str_vec_t &passive_list = str_vec_t()) const;

 private:
  std::string name_;
  str_vec_t keys_;
  std::vector<EFlt3DArray> arrays_;
};

// ------------------------------------------------------- fluid properties

// The slice of EnzoPhysicsFluidProps the method reads: ideal-gas gamma
// (EnzoEOSIdeal) and the density / pressure floors
// (EnzoFluidFloorConfig); the dual-energy formalism is not provided.
class EnzoPhysicsFluidProps {
 public:
  EnzoPhysicsFluidProps(double gamma, double density_floor,
                        double pressure_floor);
  double gamma() const { return gamma_; }
  double density_floor() const { return density_floor_; }
  double pressure_floor() const { return pressure_floor_; }
  // Physics:fluid_props:dual_energy (EnzoConfig.cpp:952-1040): "disabled",
  // "modern" (one eta, default 0.001) or "bryan95" (eta_1, eta_2)
  void set_dual_energy(const std::string &type, const std::vector<double> &eta);
  bool dual_energy_enabled() const { return dual_type_ != "disabled"; }
  bool dual_energy_modern() const { return dual_type_ == "modern"; }
  double dual_energy_eta() const { return eta_[0]; }

  // EnzoPhysicsFluidProps::primitive_from_integration (:64-138)
  void primitive_from_integration(const EnzoEFltArrayMap &integration_map,
                                  EnzoEFltArrayMap &primitive_map,
                                  int stale_depth, const str_vec_t &passive_list,
                                  bool ignore_grackle = false) const;
  // EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync (:158-290), in place
  void apply_floor_to_energy_and_sync(EnzoEFltArrayMap &integration_map,
                                      const int stale_depth) const;
  // EnzoPhysicsFluidProps::pressure_from_integration (:142-156)
  void pressure_from_integration(const EnzoEFltArrayMap &integration_map,
                                 const EFlt3DArray &pressure, int stale_depth,
                                 bool ignore_grackle = false) const;

 private:
  double gamma_, density_floor_, pressure_floor_;
  std::string dual_type_ = "disabled";
  double eta_[2] = {0.0, 0.0};
};

namespace enzo {
// the process-wide fluid properties (Physics "fluid_props")
EnzoPhysicsFluidProps *fluid_props();
void set_fluid_props(const EnzoPhysicsFluidProps &props);
}  // namespace enzo

// ---------------------------------------------------------- Cello slices

// "key = value" parameters of one Method group (Cello ParameterGroup);
// values are given as text, separated by ';' or newlines.
class ParameterGroup {
 public:
  explicit ParameterGroup(const std::string &text = "",
                          std::string group_path = "Method:mhd_vlct");
  void set(const std::string &name, const std::string &value);
  const std::string *param(const std::string &name) const;
  std::string value_string(const std::string &name, const std::string &dflt) const;
  double value_float(const std::string &name, double dflt) const;
  std::string full_name(const std::string &name) const;
  const std::string &get_group_path() const { return path_; }

 private:
  std::string path_;
  std::map<std::string, std::string> values_;
};

// Cello's field groups (cello::field_groups(), src/Cello/data_Grouping.hpp):
// group name -> ordered field list.  It starts with the groups of the
// reference's flux-correction decks (input/FluxCorrect/
// inclined_contact.incl): "conserved" = density, total_energy, velocity_x,
// velocity_y, velocity_z; "make_field_conservative" = velocity_x,
// velocity_y, velocity_z, total_energy.
class Grouping {
 public:
  Grouping();
  void add(const std::string &field, const std::string &group);
  void clear(const std::string &group);
  bool is_in(const std::string &field, const std::string &group) const;
  str_vec_t group_list(const std::string &group) const;
  int size(const std::string &group) const { return (int)group_list(group).size(); }

 private:
  std::map<std::string, str_vec_t> groups_;
};

namespace cello {
Grouping *field_groups();
}

// Cello FluxData (src/Cello/data_FluxData.hpp) with single-array storage
// (data_FluxData.cpp:14-110) in device memory: for each listed field, axis
// and face, the block's face fluxes and its neighbours' (the layout of
// include/enzoe_vlct.h).  allocate() zeroes it, as the reference's
// std::vector does.
class FluxData {
 public:
  void allocate(int nx, int ny, int nz, const str_vec_t &field_list);
  void deallocate();
  bool allocated() const { return allocated_; }
  int num_fields() const { return (int)fields_.size(); }
  const std::string &field_name(int i_f) const { return fields_[i_f]; }
  const str_vec_t &field_list() const { return fields_; }
  // FaceFluxes::flux_array of block_fluxes / neighbor_fluxes (device)
  double *block_fluxes(int axis, int face, int i_f) const;
  double *neighbor_fluxes(int axis, int face, int i_f) const;
  double *flux_array() const { return static_cast<double *>(storage_.get()); }
  long long num_values() const { return size_; }
  const std::array<int, 3> &size() const { return n_; }

 private:
  long long offset_(int axis, int face, int i_f, int nbr) const;
  std::shared_ptr<void> storage_;  // kept by deallocate() for the next cycle
  bool allocated_ = false;
  str_vec_t fields_;
  std::array<int, 3> n_{{0, 0, 0}};
  long long size_ = 0;
};

// The slice of a Cello Block the method touches: its permanent fields
// (device), cell widths, dt, the compute_done continuation, its refinement
// level and its neighbours' (Block::level, Block::face_level) and its
// FluxData.
class Block {
 public:
  // device-resident fields, one allocation holding every permanent field
  // (density, velocity_x..z, total_energy, pressure and for MHD
  // bfield_x..z, bfieldi_x..z), ghost zones included
  Block(const std::array<int, 3> &active_xyz, int ghost_depth,
        const std::array<double, 3> &cell_width, bool mhd);
  // fields owned by the caller
  explicit Block(const enzoe_block &fields);
  // a passively advected scalar (a density, cell-centred, the block's
  // shape), owned by the caller: the "color" field group of the reference
  void add_passive_scalar(const std::string &name, double *data);
  // another cell-centred field of the block's shape owned by the caller
  // (e.g. acceleration_x/y/z, which make EnzoMethodMHDVlct add gravity)
  void add_field(const std::string &name, double *data);
  const str_vec_t &passive_scalars() const { return passive_; }

  EFlt3DArray field_view(const std::string &name) const;
  bool has_field(const std::string &name) const;
  const double *CellWidth() const { return cell_width_.data(); }
  int ghost_depth() const { return ghost_; }
  double dt() const { return dt_; }
  void set_dt(double dt) { dt_ = dt; }
  bool is_leaf() const { return true; }
  void compute_done() { ++compute_done_count_; }
  int compute_done_count() const { return compute_done_count_; }
  // the fields as the library's block descriptor
  enzoe_block desc() const;
  // active cells along x, y, z (Field::size)
  std::array<int, 3> active_size() const;
  int level() const { return level_; }
  void set_level(int level) { level_ = level; }
  // level of the neighbour across (axis, face); the block's own by default
  int face_level(int axis, int face) const;
  void set_face_level(int axis, int face, int level);
  FluxData *flux_data() { return &flux_data_; }

 private:
  void bind_(const enzoe_block &b);
  std::shared_ptr<void> storage_;
  std::map<std::string, EFlt3DArray> fields_;
  str_vec_t passive_;
  std::array<double, 3> cell_width_{{0., 0., 0.}};
  int ghost_ = 0;
  double dt_ = 0.;
  int compute_done_count_ = 0;
  int level_ = 0;
  std::array<int, 6> face_level_{{-1, -1, -1, -1, -1, -1}};
  FluxData flux_data_;
};

// Block::refresh_load_flux_face_ toward a coarser neighbour
// (src/Cello/control_refresh.cpp:1458-1500) and the message's arrival
// (data_DataMsg.cpp:383-396): the fine block's fluxes on (axis, face),
// coarsened into quadrant `child` (its child index in its parent), summed
// into the coarse block's neighbour fluxes on (axis, 1 - face).
void refresh_flux_face_to_coarse(Block *fine, Block *coarse, int axis, int face,
                                 const std::array<int, 3> &child);

// Cello Method: compute(Block*) + timestep(Block*) + courant
// (src/Cello/problem_Method.hpp:14-127).
class Method {
 public:
  virtual ~Method() = default;
  virtual void compute(Block *block) throw() = 0;
  virtual std::string name() throw() = 0;
  virtual double timestep(Block *block) throw() = 0;
  double courant() const { return courant_; }
  void set_courant(double courant) { courant_ = courant; }

 protected:
  double courant_ = 1.0;
};

// --------------------------------------------------------- reconstruction

class EnzoReconstructor {
 public:
  // EnzoReconstructor::construct_reconstructor (EnzoReconstructor.cpp:14-44)
  static EnzoReconstructor *construct_reconstructor(
      const str_vec_t &active_primitive_keys, std::string name,
      enzo_float theta_limiter);

  explicit EnzoReconstructor(str_vec_t active_primitive_keys)
      : active_key_names_(std::move(active_primitive_keys)) {}
  virtual ~EnzoReconstructor() = default;

  // EnzoReconstructor.hpp:61-64
  virtual void reconstruct_interface(const EnzoEFltArrayMap &prim_map,
                                     EnzoEFltArrayMap &priml_map,
                                     EnzoEFltArrayMap &primr_map, const int dim,
                                     const int stale_depth,
                                     const str_vec_t &passive_list);
  virtual int total_staling_rate() = 0;
  virtual int immediate_staling_rate() = 0;

 protected:
  virtual const char *device_name_() const = 0;
  virtual enzo_float theta_() const { return 1.5; }
  str_vec_t active_key_names_;
};

class EnzoReconstructorNN : public EnzoReconstructor {
 public:
  explicit EnzoReconstructorNN(const str_vec_t &active_primitive_keys)
      : EnzoReconstructor(active_primitive_keys) {}
  // EnzoReconstructorNN.hpp:31-35
  int total_staling_rate() override { return 1; }
  int immediate_staling_rate() override { return 0; }

 protected:
  const char *device_name_() const override { return "nn"; }
};

// the two slope limiters of EnzoReconstructorPLM.hpp:257-346
struct PLM_EnzoRKLimiter {
  static constexpr const char *device_name = "plm";
};
struct PLM_AthenaLimiter {
  static constexpr const char *device_name = "plm_athena";
};

template <class Limiter>
class EnzoReconstructorPLM : public EnzoReconstructor {
 public:
  EnzoReconstructorPLM(const str_vec_t &active_primitive_keys,
                       enzo_float theta_limiter)
      : EnzoReconstructor(active_primitive_keys), theta_limiter_(theta_limiter) {}
  // EnzoReconstructorPLM.hpp:153-157
  int total_staling_rate() override { return 2; }
  int immediate_staling_rate() override { return 1; }

 protected:
  const char *device_name_() const override { return Limiter::device_name; }
  enzo_float theta_() const override { return theta_limiter_; }
  enzo_float theta_limiter_;
};

using EnzoReconstructorPLMEnzoRKLim = EnzoReconstructorPLM<PLM_EnzoRKLimiter>;
using EnzoReconstructorPLMAthenaLim = EnzoReconstructorPLM<PLM_AthenaLimiter>;

// ---------------------------------------------------------------- Riemann

class EnzoRiemann {
 public:
  // EnzoRiemann.hpp:44-60
  struct FactoryArgs {
    std::string solver;
    bool mhd;
    bool internal_energy;
  };
  // EnzoRiemann::construct_riemann (EnzoRiemann.cpp:26-68)
  static EnzoRiemann *construct_riemann(const FactoryArgs &factory_args) noexcept;

  // internal_energy: the dual-energy formalism adds "internal_energy" to
  // the integration quantities (EnzoRiemannImpl.hpp:240-262)
  EnzoRiemann(std::string solver, bool mhd, bool internal_energy = false);
  virtual ~EnzoRiemann() = default;

  // EnzoRiemann.hpp:125-129 (interface_velocity: dual energy only)
  void solve(const EnzoEFltArrayMap &prim_map_l,
             const EnzoEFltArrayMap &prim_map_r, EnzoEFltArrayMap &flux_map,
             const int dim, const int stale_depth,
             const str_vec_t &passive_list,
             const EFlt3DArray *const interface_velocity) const;
  // EnzoRiemann.hpp:134-141
  const str_vec_t &integration_quantity_keys() const { return integration_keys_; }
  const str_vec_t &primitive_quantity_keys() const { return primitive_keys_; }
  bool mhd() const { return mhd_; }

 private:
  std::string solver_;
  bool mhd_;
  str_vec_t integration_keys_, primitive_keys_;
};

// ------------------------------------------------- constrained transport

// EnzoBfieldMethod (toolkit/EnzoBfieldMethod.{hpp,cpp}): the partial
// timestep state machine (:32-63).
class EnzoBfieldMethod {
 public:
  explicit EnzoBfieldMethod(int num_partial_timesteps);
  virtual ~EnzoBfieldMethod() = default;
  void register_target_block(Block *block) noexcept;
  void increment_partial_timestep() noexcept;
  int num_partial_timesteps() const { return num_partial_timesteps_; }
  int partial_timestep_index() const { return partial_timestep_index_; }

  virtual void check_required_fields() const noexcept = 0;
  virtual void correct_reconstructed_bfield(EnzoEFltArrayMap &l_map,
                                            EnzoEFltArrayMap &r_map, int dim,
                                            int stale_depth) noexcept = 0;
  virtual void identify_upwind(const EnzoEFltArrayMap &flux_map, int dim,
                               int stale_depth) noexcept = 0;
  virtual void update_all_bfield_components(
      EnzoEFltArrayMap &cur_integration_map, const EnzoEFltArrayMap &xflux_map,
      const EnzoEFltArrayMap &yflux_map, const EnzoEFltArrayMap &zflux_map,
      EnzoEFltArrayMap &out_centered_bfield_map, enzo_float dt,
      int stale_depth) noexcept = 0;

 protected:
  virtual void register_target_block_(Block *target_block,
                                      bool first_initialization) noexcept = 0;
  void require_registered_block_() const noexcept;
  Block *target_block_ = nullptr;

 private:
  int num_partial_timesteps_;
  int partial_timestep_index_ = -1;
};

class EnzoBfieldMethodCT : public EnzoBfieldMethod {
 public:
  explicit EnzoBfieldMethodCT(int num_partial_timesteps);

  void check_required_fields() const noexcept override {}
  void correct_reconstructed_bfield(EnzoEFltArrayMap &l_map,
                                    EnzoEFltArrayMap &r_map, int dim,
                                    int stale_depth) noexcept override;
  void identify_upwind(const EnzoEFltArrayMap &flux_map, int dim,
                       int stale_depth) noexcept override;
  void update_all_bfield_components(EnzoEFltArrayMap &cur_integration_map,
                                    const EnzoEFltArrayMap &xflux_map,
                                    const EnzoEFltArrayMap &yflux_map,
                                    const EnzoEFltArrayMap &zflux_map,
                                    EnzoEFltArrayMap &out_centered_bfield_map,
                                    enzo_float dt,
                                    int stale_depth) noexcept override;

  // static physics methods of EnzoBfieldMethodCT.hpp
  static void compute_center_efield(const int dim, EFlt3DArray &efield,
                                    const EnzoEFltArrayMap &integration_map,
                                    int stale_depth = 0);
  static void compute_edge_efield(int dim, const EFlt3DArray &center_efield,
                                  const EFlt3DArray &edge_efield,
                                  const EnzoEFltArrayMap &jflux_map,
                                  const EnzoEFltArrayMap &kflux_map,
                                  const std::array<EFlt3DArray, 3> &weight_l,
                                  int stale_depth = 0);
  static void compute_all_edge_efields(const EnzoEFltArrayMap &integration_map,
                                       const EnzoEFltArrayMap &xflux_map,
                                       const EnzoEFltArrayMap &yflux_map,
                                       const EnzoEFltArrayMap &zflux_map,
                                       EFlt3DArray &center_efield,
                                       std::array<EFlt3DArray, 3> &edge_efield_l,
                                       const std::array<EFlt3DArray, 3> &weight_l,
                                       int stale_depth = 0);
  static void update_bfield(const enzo_float *cell_widths, int dim,
                            const std::array<EFlt3DArray, 3> &efield_l,
                            const EFlt3DArray &cur_interface_bfield,
                            const EFlt3DArray &out_interface_bfield,
                            enzo_float dt, int stale_depth = 0);
  static void compute_center_bfield(int dim, const EFlt3DArray &bfieldc_comp,
                                    const EFlt3DArray &bfieldi_comp,
                                    int stale_depth = 0);

 protected:
  void register_target_block_(Block *target_block,
                              bool first_initialization) noexcept override;

 private:
  std::array<EFlt3DArray, 3> bfieldi_l_, temp_bfieldi_l_, weight_l_,
      edge_efield_l_;
  EFlt3DArray center_efield_;
  const enzo_float *cell_widths_ = nullptr;
};

// ---------------------------------------------------- quantity updating

class EnzoIntegrationQuanUpdate {
 public:
  // EnzoIntegrationQuanUpdate.cpp:61-79
  EnzoIntegrationQuanUpdate(const str_vec_t &integration_quantity_keys,
                            const bool skip_B_update) throw();
  void clear_dUcons_map(EnzoEFltArrayMap &dUcons_map, enzo_float value,
                        const str_vec_t &passive_list) const noexcept;
  void accumulate_flux_component(int dim, double dt, enzo_float cell_width,
                                 const EnzoEFltArrayMap &flux_map,
                                 EnzoEFltArrayMap &dUcons_map, int stale_depth,
                                 const str_vec_t &passive_list) const noexcept;
  void update_quantities(EnzoEFltArrayMap &initial_integration_map,
                         const EnzoEFltArrayMap &dUcons_map,
                         EnzoEFltArrayMap &out_integration_map,
                         const int stale_depth,
                         const str_vec_t &passive_list) const;
  const str_vec_t &integration_keys() const { return integration_keys_; }

 private:
  str_vec_t integration_keys_;
};

// ------------------------------------------------------- stage commands

// EnzoMHDIntegratorStageCommands.hpp: arguments of the integrator
struct EnzoMHDIntegratorStageArgPack {
  std::string rsolver;
  str_vec_t recon_names;
  double theta_limiter;
  std::string mhd_choice;
};

// EnzoSourceInternalEnergy (toolkit/EnzoSourceInternalEnergy.hpp,
// .cpp:16-92): the internal-energy source term -p div(v) of the dual-energy
// formalism along one dim, from the interface velocities of the Riemann
// solve; accumulated into dUcons_map's "internal_energy"
class EnzoSourceInternalEnergy {
 public:
  void calculate_source(const int dim, const double dt, const enzo_float cell_width,
                        const EnzoEFltArrayMap &prim_map, EnzoEFltArrayMap &dUcons_map,
                        const EFlt3DArray &interface_velocity,
                        const int stale_depth) const throw();
};

// EnzoSourceGravity (toolkit/EnzoSourceGravity.hpp, .cpp:17-69)
class EnzoSourceGravity {
 public:
  // dUcons momentum += dt rho a and energy += dt rho (v . a) over
  // [stale_depth, m - stale_depth); accel_map holds acceleration_x/y/z
  void calculate_source(const double dt, const EnzoEFltArrayMap &prim_map,
                        EnzoEFltArrayMap &dUcons_map, const EnzoEFltArrayMap &accel_map,
                        const int stale_depth) const noexcept;
};

class EnzoMHDIntegratorStageCommands {
 public:
  enum class bfield_choice { no_bfield, unsafe_const_uniform, constrained_transport };

  explicit EnzoMHDIntegratorStageCommands(const EnzoMHDIntegratorStageArgPack &args);
  ~EnzoMHDIntegratorStageCommands();
  EnzoMHDIntegratorStageCommands(const EnzoMHDIntegratorStageCommands &) = delete;
  EnzoMHDIntegratorStageCommands &operator=(const EnzoMHDIntegratorStageCommands &) = delete;

  // EnzoMHDIntegratorStageCommands.cpp:102-203
  void compute_update_stage(EnzoEFltArrayMap tstep_begin_integration_map,
                            EnzoEFltArrayMap cur_stage_integration_map,
                            EnzoEFltArrayMap out_integration_map,
                            EnzoEFltArrayMap primitive_map,
                            EnzoEFltArrayMap priml_map,
                            EnzoEFltArrayMap primr_map,
                            std::array<EnzoEFltArrayMap, 3> flux_maps_xyz,
                            EnzoEFltArrayMap dUcons_map,
                            const EnzoEFltArrayMap &accel_map,
                            EFlt3DArray interface_vel_arr,
                            const str_vec_t &passive_list,
                            EnzoBfieldMethod *bfield_method,
                            unsigned short stage_index, double cur_dt,
                            int stale_depth,
                            const std::array<enzo_float, 3> cell_widths_xyz) const noexcept;
  // EnzoMHDIntegratorStageCommands.cpp:299-366 (no courant factor)
  double timestep(EnzoEFltArrayMap integration_map, const EFlt3DArray &pressure,
                  const double dx, const double dy, const double dz) const noexcept;

  int staling_from_stage(int stage_index) const noexcept;
  EnzoBfieldMethod *construct_bfield_method(int num_partial_timesteps) const noexcept;
  str_vec_t integration_quantity_keys() const noexcept;
  str_vec_t primitive_quantity_keys() const noexcept;
  str_vec_t dUcons_map_keys() const noexcept;
  bool is_pure_hydro() const noexcept { return mhd_choice_ == bfield_choice::no_bfield; }

 private:
  static bfield_choice parse_bfield_choice_(std::string choice) noexcept;
  void check_valid_stage_index_(int stage_index) const noexcept;
  void compute_flux_(const int dim, const double cur_dt,
                     const enzo_float cell_width, EnzoEFltArrayMap &primitive_map,
                     EnzoEFltArrayMap &priml_map, EnzoEFltArrayMap &primr_map,
                     EnzoEFltArrayMap &flux_map, EnzoEFltArrayMap &dUcons_map,
                     const EFlt3DArray *const interface_velocity_arr_ptr,
                     EnzoReconstructor &reconstructor,
                     EnzoBfieldMethod *bfield_method, const int stale_depth,
                     const str_vec_t &passive_list) const noexcept;

  EnzoRiemann *riemann_solver_ = nullptr;
  std::vector<std::unique_ptr<EnzoReconstructor>> reconstructors_;
  EnzoIntegrationQuanUpdate *integration_quan_updater_ = nullptr;
  bfield_choice mhd_choice_;
};

// EnzoVlctScratchSpace (EnzoMethodMHDVlct.hpp:244-273)
struct EnzoVlctScratchSpace {
  EnzoVlctScratchSpace(const std::array<int, 3> &field_shape,
                       const str_vec_t &integration_keys,
                       const str_vec_t &primitive_keys,
                       const str_vec_t &dUcons_keys,
                       const str_vec_t &passive_list, bool dual_energy = false);
  EnzoEFltArrayMap temp_integration_map, primitive_map, priml_map, primr_map,
      xflux_map, yflux_map, zflux_map, dUcons_map;
  // the interface velocities of the dual-energy formalism (cell-shaped,
  // sliced per dim like the reconstructed states), null otherwise
  EFlt3DArray interface_vel_arr;
  str_vec_t passive_list;
};

// ----------------------------------------------------------------- method

class EnzoMethodMHDVlct : public Method {
 public:
  // EnzoMethodMHDVlct.cpp:89-150.  Parameters (defaults as the reference):
  // mhd_choice (required), riemann_solver ("hlld"), reconstruct_method
  // ("plm"), theta_limiter (1.5), time_scheme ("vl"), courant (0.3 for
  // vl, 1.0 for euler).  Device extensions: "device_fused" (true: one fused
  // library launch sequence per compute; false: the stage-by-stage plugin
  // composition below) and "precision" ("double" or "mixed", fused only).
  // store_fluxes_for_corrections (hydro only, as the reference asserts):
  // each compute allocates the block's FluxData for the "conserved" group
  // and fills it with the final stage's face fluxes.
  EnzoMethodMHDVlct(ParameterGroup p, bool store_fluxes_for_corrections);
  ~EnzoMethodMHDVlct() override;

  // EnzoMethodMHDVlct.cpp:356-500; calls block->compute_done() once
  void compute(Block *block) throw() override;
  // several blocks sharing dt in one batched launch; compute_done is called
  // on each block once the launch sequence is enqueued
  void compute_blocks(const std::vector<Block *> &blocks) throw();
  std::string name() throw() override { return "mhd_vlct"; }
  // EnzoMethodMHDVlct.cpp:551-588
  double timestep(Block *block) throw() override;

  bool device_fused() const { return fused_; }

 private:
  EnzoEFltArrayMap get_integration_map_(Block *block,
                                        const str_vec_t &passive_list) const noexcept;
  // EnzoMethodMHDVlct.cpp:219-231: acceleration_x/y/z or an empty map
  static EnzoEFltArrayMap get_accel_map_(Block *block) noexcept;
  EnzoVlctScratchSpace *get_scratch_ptr_(const std::array<int, 3> &field_shape,
                                         const str_vec_t &passive_list) noexcept;
  void compute_staged_(Block *block) noexcept;
  // EnzoMethodMHDVlct.cpp:333-355
  void allocate_FC_flux_buffer_(Block *block) const;
  std::vector<int> flux_slots_(const FluxData &fd) const;

  bool store_fluxes_for_corrections_ = false;
  std::string time_scheme_;
  EnzoMHDIntegratorStageArgPack *integrator_arg_pack_ = nullptr;
  EnzoMHDIntegratorStageCommands *integrator_ = nullptr;
  EnzoVlctScratchSpace *scratch_space_ = nullptr;
  EnzoBfieldMethod *bfield_method_ = nullptr;
  str_vec_t integration_field_list_, primitive_field_list_;
  bool fused_ = true;
  int precision_ = ENZOE_PREC_DOUBLE;
  enzoe_vlct_t device_ = nullptr;
  // debug_checks = true: the fused step counts non-finite updates and
  // applied floors on the device; a nonzero count ends the run as the
  // reference's debug builds do (switched on at the first compute: the
  // constructor makes no device call)
  bool debug_checks_ = false;
  bool debug_on_ = false;
  void debug_check_(const char *where);
};

// -------------------------------------------------------- initial values

// EnzoInitialInclinedWave (src/Enzo/initial/EnzoInitialInclinedWave.cpp):
// a linear wave along the direction rotated by (alpha, beta).  Parameters
// (defaults of :647-666): alpha (0), beta (0), amplitude (1e-6), lambda (1),
// positive_vel (true), wave_type (required: "fast", "alfven", "slow",
// "mhd_entropy", "sound", "hd_entropy", "hd_transv_entropy_v1",
// "hd_transv_entropy_v2").  The values are computed on the host, as the
// reference's initializers do, and copied into the block's device fields
// (ghost zones included); "circ_alfven", "jeans" (gravity) and
// parallel_vel are not provided (ERROR).
class EnzoInitialInclinedWave {
 public:
  explicit EnzoInitialInclinedWave(ParameterGroup p);
  // enforce_block (:722-772) for a leaf block whose active region starts at
  // `lower` (the block's lower corner in the domain)
  void enforce_block(Block *block, const std::array<double, 3> &lower) const;

 private:
  double alpha_, beta_, amplitude_, lambda_;
  bool pos_vel_;
  std::string wave_type_;
};

// EnzoInitialShockTube (src/Enzo/initial/EnzoInitialShockTube.cpp): the
// "rj2a" (Ryu & Jones 1995, fig. 2a) and "sod" tubes along aligned_ax
// ("x", "y", "z"), with axis_velocity, transverse_velocity and
// flip_initialize.  Host values copied into the block's device fields.
class EnzoInitialShockTube {
 public:
  explicit EnzoInitialShockTube(ParameterGroup p);
  // enforce_block (:183-292) for a block whose active region is
  // [lower, upper)
  void enforce_block(Block *block, const std::array<double, 3> &lower,
                     const std::array<double, 3> &upper) const;

 private:
  std::string setup_name_;
  int aligned_ax_ = 0;
  double axis_velocity_ = 0.0, trans_velocity_ = 0.0;
  bool flipped_initialize_ = false;
};

// ------------------------------------------------ coarse/fine transfers

// Cello's precision_enum (src/Cello/cello.hpp:152-163)
enum precision_enum {
  precision_unknown, precision_default, precision_single, precision_double,
  precision_extended64, precision_extended80, precision_extended96, precision_quadruple
};
typedef int precision_type;

// Cello Prolong (src/Cello/problem_Prolong.hpp): values_f / values_c are
// device arrays, dims (x, y, z); only double precision (enzo_float) is
// provided.  Launches on the stream of enzoe_host_set_stream.
class Prolong {
 public:
  virtual ~Prolong() = default;
  virtual void apply(precision_type precision, void *values_f, int m3_f[3], int o3_f[3],
                     int n3_f[3], const void *values_c, int m3_c[3], int o3_c[3],
                     int n3_c[3], bool accumulate = false) = 0;
  virtual std::string name() const = 0;
  // cells of padding needed around the coarse region (problem_Prolong.hpp:61)
  virtual int coarse_padding() const { return 0; }
};

// problem_ProlongLinear.cpp:25-370
class ProlongLinear : public Prolong {
 public:
  void apply(precision_type precision, void *values_f, int m3_f[3], int o3_f[3],
             int n3_f[3], const void *values_c, int m3_c[3], int o3_c[3], int n3_c[3],
             bool accumulate = false) override;
  std::string name() const override { return "linear"; }
};

// src/Enzo/mesh/EnzoProlong.cpp: types "3A", "2A", "2B"; with use_linear
// it is ProlongLinear on the coarse region shrunk by one cell per side, for
// accumulate ProlongLinear as given (:49-97); otherwise Enzo's
// interpolate() (apply_, :112-190) on the device (enzoe_enzo_prolong,
// rank 3).
class EnzoProlong : public Prolong {
 public:
  EnzoProlong(std::string type, bool positive, bool use_linear);
  void apply(precision_type precision, void *values_f, int m3_f[3], int o3_f[3],
             int n3_f[3], const void *values_c, int m3_c[3], int o3_c[3], int n3_c[3],
             bool accumulate = false) override;
  std::string name() const override { return "enzo"; }
  int coarse_padding() const override { return 1; }

 private:
  int method_ = -1;
  int positive_ = 0;
  bool use_linear_ = false;
};

// Cello Restrict (src/Cello/problem_Restrict.hpp); returns bytes written.
class Restrict {
 public:
  virtual ~Restrict() = default;
  virtual int apply(precision_type precision, void *values_c, int nd3_c[3], int im3_c[3],
                    int n3_c[3], const void *values_f, int nd3_f[3], int im3_f[3],
                    int n3_f[3], bool accumulate = false) = 0;
  virtual std::string name() const = 0;
};

// problem_RestrictLinear.cpp:20-198
class RestrictLinear : public Restrict {
 public:
  int apply(precision_type precision, void *values_c, int nd3_c[3], int im3_c[3],
            int n3_c[3], const void *values_f, int nd3_f[3], int im3_f[3], int n3_f[3],
            bool accumulate = false) override;
  std::string name() const override { return "linear"; }
};

// src/Enzo/mesh/EnzoRestrict.cpp: writes nothing and returns 0 (:36-45)
class EnzoRestrict : public Restrict {
 public:
  explicit EnzoRestrict(std::string restrict_type = "enzo") { (void)restrict_type; }
  int apply(precision_type, void *, int[3], int[3], int[3], const void *, int[3], int[3],
            int[3], bool = false) override {
    return 0;
  }
  std::string name() const override { return "enzo"; }
};

// ------------------------------------------------------- flux correction

// Cello MethodFluxCorrect (src/Cello/problem_MethodFluxCorrect.cpp): on
// every leaf block, replaces the fluxes through faces whose neighbour is
// finer by the neighbour's (coarsened) fluxes, then sums the conserved
// fields over the blocks and reports how many digits of the first cycle's
// sums are kept.
class MethodFluxCorrect : public Method {
 public:
  // Parameters: group ("conserved"), enable (true), min_digits (a float
  // for density, or a list "[field, digits, ...]"), :57-63, 15-52.
  explicit MethodFluxCorrect(ParameterGroup p);
  MethodFluxCorrect(std::string group, bool enable,
                    std::map<std::string, double> min_digits_map);

  // compute() of one block and the sum over it alone
  void compute(Block *block) throw() override;
  // every leaf block of a cycle: corrections, the sum reduction over the
  // blocks (compute_continue_refresh / compute_continue_sum_fields,
  // :134-262), FluxData deallocated, compute_done on each block
  void compute_blocks(const std::vector<Block *> &blocks) throw();
  std::string name() throw() override { return "flux_correct"; }
  double timestep(Block *) throw() override { return 1.7976931348623157e308; }

  // the last reduction: sums of the group's fields (times density for
  // "make_field_conservative" ones, scaled by 2^-(3 level)), the first
  // cycle's, and -log10 of their relative difference
  const std::vector<long double> &field_sums() const { return field_sum_; }
  double conserved_digits(const std::string &field) const;
  // false if a field of min_digits kept fewer digits (the unit test of the
  // reference, :248-253)
  bool min_digits_ok() const { return min_digits_ok_; }

  // :425-560 (public for plugin-level tests)
  void flux_correct_(Block *block);

 private:
  void local_sums_(Block *block, std::vector<long double> &reduce) const;

  std::string group_;
  bool enable_ = true;
  std::map<std::string, double> min_digits_map_;
  str_vec_t fields_;
  std::vector<long double> field_sum_, field_sum_0_;
  std::map<std::string, double> digits_;
  bool min_digits_ok_ = true;
  int cycle_ = 0;
};

// enzoe_host.cpp -- the reference's VL+CT plugin classes over the gfx950
// library.  Control flow, names, defaults and error messages follow the
// reference files cited at each function; the arithmetic is all on the
// device, one C-ABI call per reference plugin call.
#include "enzoe_host.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

// ---------------------------------------------------------------- errors

void enzoe_host_error(const char *function, const std::string &message) {
  std::fprintf(stderr, "ERROR in %s: %s\n", function, message.c_str());
  std::fflush(stderr);
  std::exit(1);
}

namespace {

thread_local enzoe_stream_t g_stream = nullptr;

const char *status_text(int rc) {
  switch (rc) {
    case ENZOE_ERR_CONFIG: return "invalid configuration";
    case ENZOE_ERR_ARG: return "invalid argument (view shape or missing key)";
    case ENZOE_ERR_HIP: return "HIP runtime error";
    case ENZOE_ERR_ALLOC: return "device allocation failed";
    case ENZOE_ERR_STATE: return "invalid state";
    case ENZOE_ERR_SIZE: return "input larger than the fixed workspace";
    default: return "unknown status";
  }
}

void check_rc(int rc, const char *where) {
  if (rc != ENZOE_OK) {
    ENZOE_ERROR(where, std::string("device library call failed: ") +
                           status_text(rc) + " (status " + std::to_string(rc) + ")");
  }
}

std::string lowered(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(),
                 [](unsigned char c) { return (char)std::tolower(c); });
  return s;
}

std::shared_ptr<void> device_alloc(std::size_t nbytes, const char *where) {
  void *p = nullptr;
  if (hipMalloc(&p, std::max<std::size_t>(nbytes, 8)) != hipSuccess)
    ENZOE_ERROR(where, "hipMalloc of " + std::to_string(nbytes) + " bytes failed");
  if (hipMemset(p, 0, std::max<std::size_t>(nbytes, 8)) != hipSuccess)
    ENZOE_ERROR(where, "hipMemset failed");
  return std::shared_ptr<void>(p, [](void *q) { (void)hipFree(q); });
}

// slot of a key in the library's maps (ENZOE_SLOT_*)
int key_slot(const std::string &key) {
  static const std::map<std::string, int> slots = {
      {"density", ENZOE_SLOT_DENSITY},       {"velocity_x", ENZOE_SLOT_VELOCITY_X},
      {"velocity_y", ENZOE_SLOT_VELOCITY_Y}, {"velocity_z", ENZOE_SLOT_VELOCITY_Z},
      {"total_energy", ENZOE_SLOT_ENERGY},   {"pressure", ENZOE_SLOT_ENERGY},
      {"bfield_x", ENZOE_SLOT_BFIELD_X},     {"bfield_y", ENZOE_SLOT_BFIELD_Y},
      {"bfield_z", ENZOE_SLOT_BFIELD_Z},     {"internal_energy", ENZOE_SLOT_INTERNAL_ENERGY}};
  auto it = slots.find(key);
  return it == slots.end() ? -1 : it->second;
}


// map restricted to `keys` (in that order)
EnzoEFltArrayMap select_keys(const EnzoEFltArrayMap &map, const str_vec_t &keys) {
  std::vector<EFlt3DArray> arrays;
  arrays.reserve(keys.size());
  for (const std::string &k : keys) arrays.push_back(map.at(k));
  return EnzoEFltArrayMap(map.name(), keys, arrays);
}

str_vec_t concat(const str_vec_t &a, const str_vec_t &b) {
  str_vec_t out(a);
  out.insert(out.end(), b.begin(), b.end());
  return out;
}

const char *const kBNames[3] = {"bfield_x", "bfield_y", "bfield_z"};
const char *const kBiNames[3] = {"bfieldi_x", "bfieldi_y", "bfieldi_z"};

}  // namespace

void enzoe_host_set_stream(enzoe_stream_t stream) { g_stream = stream; }
enzoe_stream_t enzoe_host_stream() { return g_stream; }

// ------------------------------------------------------------------ views

void CSlice::resolve(int n, int *lo, int *hi) const {
  int a = has_start_ ? (start_ < 0 ? n + start_ : start_) : 0;
  int b = has_stop_ ? (stop_ < 0 ? n + stop_ : stop_) : n;
  ENZOE_ASSERT("CSlice", "slice outside the axis or empty",
               a >= 0 && b <= n && a < b);
  *lo = a;
  *hi = b;
}

EFlt3DArray::EFlt3DArray(int mz, int my, int mx) {
  ENZOE_ASSERT("EFlt3DArray", "negative extent", mz >= 0 && my >= 0 && mx >= 0);
  owner_ = device_alloc((std::size_t)mz * my * mx * sizeof(double), "EFlt3DArray");
  data_ = static_cast<double *>(owner_.get());
  shape_ = {{mz, my, mx}};
  stride_ = {{(long long)my * mx, (long long)mx, 1}};
}

EFlt3DArray::EFlt3DArray(double *data, int mz, int my, int mx)
    : data_(data), shape_{{mz, my, mx}},
      stride_{{(long long)my * mx, (long long)mx, 1}} {}

EFlt3DArray::EFlt3DArray(std::shared_ptr<void> owner, double *data,
                         const std::array<int, 3> &shape,
                         const std::array<long long, 3> &stride)
    : owner_(std::move(owner)), data_(data), shape_(shape), stride_(stride) {}

EFlt3DArray EFlt3DArray::subarray(const CSlice &z, const CSlice &y,
                                  const CSlice &x) const {
  const CSlice *s[3] = {&z, &y, &x};
  std::array<int, 3> shape;
  long long offset = 0;
  for (int a = 0; a < 3; a++) {
    int lo, hi;
    s[a]->resolve(shape_[a], &lo, &hi);
    shape[a] = hi - lo;
    offset += lo * stride_[a];
  }
  return EFlt3DArray(owner_, data_ + offset, shape, stride_);
}

enzoe_farray EFlt3DArray::desc() const {
  enzoe_farray d;
  d.data = data_;
  for (int a = 0; a < 3; a++) {
    d.shape[a] = shape_[a];
    d.stride[a] = stride_[a];
  }
  return d;
}

EnzoPermutedCoordinates::EnzoPermutedCoordinates(int i_axis) : i_(i_axis) {
  ENZOE_ASSERT("EnzoPermutedCoordinates", "i_axis must be 0, 1, or 2",
               i_axis >= 0 && i_axis <= 2);
}

EFlt3DArray EnzoPermutedCoordinates::get_subarray(const EFlt3DArray &a,
                                                  const CSlice &k_slice,
                                                  const CSlice &j_slice,
                                                  const CSlice &i_slice) const {
  // axis index (0 = x) -> slice; arrays are indexed (z, y, x)
  CSlice by_axis[3];
  by_axis[i_axis()] = i_slice;
  by_axis[j_axis()] = j_slice;
  by_axis[k_axis()] = k_slice;
  return a.subarray(by_axis[2], by_axis[1], by_axis[0]);
}

EFlt3DArray EnzoPermutedCoordinates::left_edge_offset(const EFlt3DArray &a,
                                                      int k_start, int j_start,
                                                      int i_start) const {
  return get_subarray(a, CSlice(k_start, nullptr), CSlice(j_start, nullptr),
                      CSlice(i_start, nullptr));
}

EnzoEFltArrayMap::EnzoEFltArrayMap(std::string name, const str_vec_t &keys,
                                   const std::array<int, 3> &shape)
    : name_(std::move(name)), keys_(keys) {
  const std::size_t per = (std::size_t)shape[0] * shape[1] * shape[2];
  std::shared_ptr<void> owner =
      device_alloc(per * keys.size() * sizeof(double), "EnzoEFltArrayMap");
  double *base = static_cast<double *>(owner.get());
  const std::array<long long, 3> stride{
      {(long long)shape[1] * shape[2], (long long)shape[2], 1}};
  for (std::size_t q = 0; q < keys.size(); q++)
    arrays_.emplace_back(owner, base + q * per, shape, stride);
}

EnzoEFltArrayMap::EnzoEFltArrayMap(std::string name, const str_vec_t &keys,
                                   const std::vector<EFlt3DArray> &arrays)
    : name_(std::move(name)), keys_(keys), arrays_(arrays) {
  ENZOE_ASSERT("EnzoEFltArrayMap", "keys and arrays differ in number",
               keys.size() == arrays.size());
  for (std::size_t q = 1; q < arrays.size(); q++)
    for (int a = 0; a < 3; a++)
      ENZOE_ASSERT("EnzoEFltArrayMap", "arrays must share one shape",
                   arrays[q].shape(a) == arrays[0].shape(a));
}

bool EnzoEFltArrayMap::contains(const std::string &key) const {
  return std::find(keys_.begin(), keys_.end(), key) != keys_.end();
}

EFlt3DArray EnzoEFltArrayMap::at(const std::string &key) const {
  auto it = std::find(keys_.begin(), keys_.end(), key);
  if (it == keys_.end())
    ENZOE_ERROR("EnzoEFltArrayMap::at", "map \"" + name_ + "\" has no key \"" + key + "\"");
  return arrays_[(std::size_t)(it - keys_.begin())];
}

EFlt3DArray EnzoEFltArrayMap::get(const std::string &key, int stale_depth) const {
  EFlt3DArray a = at(key);
  if (stale_depth == 0) return a;
  ENZOE_ASSERT("EnzoEFltArrayMap::get", "stale_depth must be non-negative",
               stale_depth > 0);
  CSlice s(stale_depth, -stale_depth);
  return a.subarray(s, s, s);
}

int EnzoEFltArrayMap::array_shape(unsigned axis) const {
  ENZOE_ASSERT("EnzoEFltArrayMap::array_shape", "empty map", !arrays_.empty());
  return arrays_[0].shape(axis);
}

EnzoEFltArrayMap EnzoEFltArrayMap::subarray_map(const CSlice &z, const CSlice &y,
                                                const CSlice &x,
                                                const std::string &name) const {
  std::vector<EFlt3DArray> arrays;
  arrays.reserve(arrays_.size());
  for (const EFlt3DArray &a : arrays_) arrays.push_back(a.subarray(z, y, x));
  return EnzoEFltArrayMap(name.empty() ? name_ : name, keys_, arrays);
}

enzoe_fmap EnzoEFltArrayMap::slot_map(const str_vec_t &passive_list) const {
  enzoe_fmap m;
  std::memset(&m, 0, sizeof(m));
  for (std::size_t q = 0; q < keys_.size(); q++) {
    int slot = key_slot(keys_[q]);
    if (slot < 0) {
      auto it = std::find(passive_list.begin(), passive_list.end(), keys_[q]);
      if (it != passive_list.end()) {
        const int idx = (int)(it - passive_list.begin());
        ENZOE_ASSERT("EnzoEFltArrayMap::slot_map",
                     "at most 8 passively advected scalars are supported",
                     idx < ENZOE_PASSIVE_SLOTS);
        slot = ENZOE_SLOT_PASSIVE0 + idx;
      }
    }
    if (slot < 0)
      ENZOE_ERROR("EnzoEFltArrayMap::slot_map",
                  "key \"" + keys_[q] + "\" has no counterpart in the device library");
    ENZOE_ASSERT("EnzoEFltArrayMap::slot_map", "two keys share one slot",
                 m.data[slot] == nullptr);
    m.data[slot] = arrays_[q].data();
    for (int a = 0; a < 3; a++) {
      if (q == 0) {
        m.shape[a] = arrays_[q].shape(a);
        m.stride[a] = arrays_[q].stride(a);
      } else {
        ENZOE_ASSERT("EnzoEFltArrayMap::slot_map",
                     "views of a map must share shape and strides",
                     m.shape[a] == arrays_[q].shape(a) &&
                         m.stride[a] == arrays_[q].stride(a));
      }
    }
  }
  return m;
}

// ------------------------------------------------------- fluid properties

EnzoPhysicsFluidProps::EnzoPhysicsFluidProps(double gamma, double density_floor,
                                             double pressure_floor)
    : gamma_(gamma), density_floor_(density_floor), pressure_floor_(pressure_floor) {
  ENZOE_ASSERT("EnzoPhysicsFluidProps", "gamma must exceed 1", gamma > 1.0);
  // EnzoMHDIntegratorStageCommands.cpp:35-40
  ENZOE_ASSERT("EnzoPhysicsFluidProps",
               "density and pressure floors must be defined",
               density_floor > 0 && pressure_floor > 0);
}

void EnzoPhysicsFluidProps::set_dual_energy(const std::string &type,
                                            const std::vector<double> &eta) {
  // parse_de_config_ (EnzoConfig.cpp:952-1010)
  const bool eta_exists = !eta.empty();
  if (type == "disabled") {
    ENZOE_ASSERT("parse_de_config_",
                 "when dual energy is disabled, \"Physics:fluid_props:dual_energy:eta\" "
                 "can't be specified", !eta_exists);
    dual_type_ = type;
  } else if (type == "modern") {
    ENZOE_ASSERT("parse_de_config_",
                 "\"Physics:fluid_props:dual_energy:eta\" must provide 1 value for the "
                 "\"modern\" dual energy formalism", eta.size() == 1 || !eta_exists);
    eta_[0] = eta_exists ? eta[0] : 0.001;
    // EnzoDualEnergyConfig::build_modern_formulation (:42-48)
    ENZOE_ASSERT("EnzoDualEnergyConfig::build_modern_formulation",
                 "eta must be non-negative", eta_[0] >= 0);
    dual_type_ = type;
  } else if (type == "bryan95") {
    ENZOE_ASSERT("parse_de_config_",
                 "\"Physics:fluid_props:dual_energy:eta\" must provide 2 values for the "
                 "\"bryan95\" dual energy formalism", eta.size() == 2 || !eta_exists);
    eta_[0] = eta_exists ? eta[0] : 0.001;
    eta_[1] = eta_exists ? eta[1] : 0.1;
    ENZOE_ASSERT("EnzoDualEnergyConfig::build_bryan95_formulation",
                 "eta1 and eta2 must be non-negative", eta_[0] >= 0 && eta_[1] >= 0);
    dual_type_ = type;
  } else {
    ENZOE_ERROR("parse_de_config_",
                "\"Physics:fluid_props:dual_energy:type\" is invalid: \"" + type + "\"");
  }
}

void EnzoPhysicsFluidProps::primitive_from_integration(
    const EnzoEFltArrayMap &integration_map, EnzoEFltArrayMap &primitive_map,
    int stale_depth, const str_vec_t &passive_list, bool) const {
  const enzoe_fmap u = integration_map.slot_map(passive_list);
  const enzoe_fmap w = primitive_map.slot_map(passive_list);
  const int mhd = u.data[ENZOE_SLOT_BFIELD_X] != nullptr;
  // pressure from the internal energy with dual energy
  // (EnzoComputePressure.cpp:127-134)
  if (dual_energy_enabled())
    check_rc(enzoe_primitive_from_integration_dual_energy(&u, &w, mhd, gamma_, stale_depth,
                                                          g_stream),
             "EnzoPhysicsFluidProps::primitive_from_integration");
  else
    check_rc(enzoe_primitive_from_integration(&u, &w, mhd, gamma_, stale_depth, g_stream),
             "EnzoPhysicsFluidProps::primitive_from_integration");
}

void EnzoPhysicsFluidProps::pressure_from_integration(
    const EnzoEFltArrayMap &integration_map, const EFlt3DArray &pressure,
    int stale_depth, bool) const {
  const enzoe_fmap u = integration_map.slot_map();
  const enzoe_farray p = pressure.desc();
  const int mhd = u.data[ENZOE_SLOT_BFIELD_X] != nullptr;
  if (dual_energy_enabled())
    check_rc(enzoe_pressure_from_integration_dual_energy(&u, &p, gamma_, stale_depth,
                                                         g_stream),
             "EnzoPhysicsFluidProps::pressure_from_integration");
  else
    check_rc(enzoe_pressure_from_integration(&u, &p, mhd, gamma_, stale_depth, g_stream),
             "EnzoPhysicsFluidProps::pressure_from_integration");
}

void EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync(EnzoEFltArrayMap &integration_map,
                                                           const int stale_depth) const {
  // EnzoPhysicsFluidProps.cpp:158-290
  const bool idual = dual_energy_enabled();
  ENZOE_ASSERT("EnzoEOSIdeal::apply_floor_to_energy_and_sync",
               "The current implementation only works when the dual energy formalism is "
               "disabled or uses the \"modern formulation\"",
               !idual || dual_energy_modern());
  str_vec_t keys = {"density", "velocity_x", "velocity_y", "velocity_z", "total_energy"};
  const bool mag = integration_map.contains("bfield_x");
  if (mag) keys.insert(keys.end(), {"bfield_x", "bfield_y", "bfield_z"});
  if (idual) keys.push_back("internal_energy");
  const enzoe_fmap u = select_keys(integration_map, keys).slot_map();
  check_rc(enzoe_apply_floor_to_energy_and_sync(&u, mag ? 1 : 0, gamma_, pressure_floor_,
                                                idual ? 1 : 0, idual ? eta_[0] : 0.0,
                                                stale_depth, g_stream),
           "EnzoPhysicsFluidProps::apply_floor_to_energy_and_sync");
}

namespace enzo {
namespace {
std::unique_ptr<EnzoPhysicsFluidProps> &props_slot() {
  static std::unique_ptr<EnzoPhysicsFluidProps> p;
  return p;
}
}  // namespace

EnzoPhysicsFluidProps *fluid_props() {
  if (!props_slot())
    ENZOE_ERROR("enzo::fluid_props", "the fluid properties were never set");
  return props_slot().get();
}

void set_fluid_props(const EnzoPhysicsFluidProps &props) {
  props_slot().reset(new EnzoPhysicsFluidProps(props));
}
}  // namespace enzo

// ---------------------------------------------------------- Cello slices

static std::string trim(const std::string &s) {
  std::size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  std::string r = s.substr(a, b - a + 1);
  if (r.size() >= 2 && r.front() == '"' && r.back() == '"') r = r.substr(1, r.size() - 2);
  return r;
}

ParameterGroup::ParameterGroup(const std::string &text, std::string group_path)
    : path_(std::move(group_path)) {
  std::string item;
  std::stringstream ss(text);
  while (std::getline(ss, item, ';')) {
    std::stringstream lines(item);
    std::string line;
    while (std::getline(lines, line)) {
      if (trim(line).empty()) continue;
      std::size_t eq = line.find('=');
      if (eq == std::string::npos)
        ENZOE_ERROR("ParameterGroup", "expected \"name = value\", got \"" + line + "\"");
      set(trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
  }
}

void ParameterGroup::set(const std::string &name, const std::string &value) {
  values_[name] = value;
}

const std::string *ParameterGroup::param(const std::string &name) const {
  auto it = values_.find(name);
  return it == values_.end() ? nullptr : &it->second;
}

std::string ParameterGroup::value_string(const std::string &name,
                                         const std::string &dflt) const {
  const std::string *v = param(name);
  return v ? *v : dflt;
}

double ParameterGroup::value_float(const std::string &name, double dflt) const {
  const std::string *v = param(name);
  if (!v) return dflt;
  char *end = nullptr;
  double x = std::strtod(v->c_str(), &end);
  if (end == v->c_str() || *end != '\0')
    ENZOE_ERROR("ParameterGroup::value_float",
                full_name(name) + " is not a number: \"" + *v + "\"");
  return x;
}

std::string ParameterGroup::full_name(const std::string &name) const {
  return path_ + ":" + name;
}

Block::Block(const std::array<int, 3> &active_xyz, int ghost_depth,
             const std::array<double, 3> &cell_width, bool mhd) {
  const int mx = active_xyz[0] + 2 * ghost_depth;
  const int my = active_xyz[1] + 2 * ghost_depth;
  const int mz = active_xyz[2] + 2 * ghost_depth;
  const std::size_t cells = (std::size_t)mx * my * mz;
  // 8 B alignment and no padding between fields, as the reference's
  // FieldData with the decks' settings (data_FieldData.cpp:319-395)
  std::size_t words = cells * (mhd ? 9 : 6);
  if (mhd) words += (std::size_t)(mx + 1) * my * mz + (std::size_t)mx * (my + 1) * mz +
                    (std::size_t)mx * my * (mz + 1);
  storage_ = device_alloc(words * sizeof(double), "Block");
  double *p = static_cast<double *>(storage_.get());
  enzoe_block b;
  std::memset(&b, 0, sizeof(b));
  b.mx = mx; b.my = my; b.mz = mz;
  b.ghost = ghost_depth;
  for (int a = 0; a < 3; a++) b.cell_width[a] = cell_width[a];
  b.density = p; p += cells;
  for (int a = 0; a < 3; a++) { b.velocity[a] = p; p += cells; }
  b.total_energy = p; p += cells;
  b.pressure = p; p += cells;
  if (mhd) {
    for (int a = 0; a < 3; a++) { b.bfield[a] = p; p += cells; }
    for (int a = 0; a < 3; a++) {
      b.bfieldi[a] = p;
      p += (std::size_t)(mx + (a == 0)) * (my + (a == 1)) * (mz + (a == 2));
    }
  }
  bind_(b);
}

Block::Block(const enzoe_block &fields) { bind_(fields); }

void Block::bind_(const enzoe_block &b) {
  ghost_ = b.ghost;
  for (int a = 0; a < 3; a++) cell_width_[a] = b.cell_width[a];
  auto put = [&](const char *name, double *p, int dz, int dy, int dx) {
    if (p) fields_[name] = EFlt3DArray(p, b.mz + dz, b.my + dy, b.mx + dx);
  };
  put("density", b.density, 0, 0, 0);
  put("velocity_x", b.velocity[0], 0, 0, 0);
  put("velocity_y", b.velocity[1], 0, 0, 0);
  put("velocity_z", b.velocity[2], 0, 0, 0);
  put("total_energy", b.total_energy, 0, 0, 0);
  put("pressure", b.pressure, 0, 0, 0);
  put("bfield_x", b.bfield[0], 0, 0, 0);
  put("bfield_y", b.bfield[1], 0, 0, 0);
  put("bfield_z", b.bfield[2], 0, 0, 0);
  put("bfieldi_x", b.bfieldi[0], 0, 0, 1);
  put("bfieldi_y", b.bfieldi[1], 0, 1, 0);
  put("bfieldi_z", b.bfieldi[2], 1, 0, 0);
  put("internal_energy", b.internal_energy, 0, 0, 0);
}

void Block::add_field(const std::string &name, double *data) {
  const EFlt3DArray &rho = field_view("density");
  ENZOE_ASSERT("Block::add_field", "\"" + name + "\" is already a field",
               fields_.count(name) == 0);
  fields_[name] = EFlt3DArray(data, rho.shape(0), rho.shape(1), rho.shape(2));
}

void Block::add_passive_scalar(const std::string &name, double *data) {
  const EFlt3DArray &rho = field_view("density");
  ENZOE_ASSERT("Block::add_passive_scalar", "\"" + name + "\" is already a field",
               fields_.count(name) == 0);
  fields_[name] = EFlt3DArray(data, rho.shape(0), rho.shape(1), rho.shape(2));
  passive_.push_back(name);
}

bool Block::has_field(const std::string &name) const {
  return fields_.count(name) != 0;
}

EFlt3DArray Block::field_view(const std::string &name) const {
  auto it = fields_.find(name);
  // EnzoMethodMHDVlct.cpp:14-20 (check_field_l_)
  if (it == fields_.end())
    ENZOE_ERROR("EnzoMethodMHDVlct", "\"" + name + "\" must be a permanent field");
  return it->second;
}

std::array<int, 3> Block::active_size() const {
  const EFlt3DArray &rho = field_view("density");
  return {rho.shape(2) - 2 * ghost_, rho.shape(1) - 2 * ghost_, rho.shape(0) - 2 * ghost_};
}

int Block::face_level(int axis, int face) const {
  const int l = face_level_[2 * axis + face];
  return l < 0 ? level_ : l;
}

void Block::set_face_level(int axis, int face, int level) {
  ENZOE_ASSERT("Block::set_face_level", "axis in [0,3) and face in [0,2)",
               axis >= 0 && axis < 3 && (face == 0 || face == 1));
  face_level_[2 * axis + face] = level;
}

void refresh_flux_face_to_coarse(Block *fine, Block *coarse, int axis, int face,
                                 const std::array<int, 3> &child) {
  FluxData *ff = fine->flux_data(), *fc = coarse->flux_data();
  ENZOE_ASSERT("Block::refresh_load_flux_face_", "both blocks hold FluxData",
               ff->allocated() && fc->allocated());
  ENZOE_ASSERT("Block::refresh_load_flux_face_", "FluxData of the same fields and size",
               ff->field_list() == fc->field_list() && ff->size() == fc->size());
  const std::array<int, 3> n = ff->size();
  check_rc(enzoe_flux_data_send_to_coarse(ff->flux_array(), fc->flux_array(),
                                          ff->num_fields(), n[0], n[1], n[2], axis, face,
                                          child.data(), g_stream),
           "Block::refresh_load_flux_face_");
}

enzoe_block Block::desc() const {
  enzoe_block b;
  std::memset(&b, 0, sizeof(b));
  const EFlt3DArray &rho = field_view("density");
  b.mz = rho.shape(0); b.my = rho.shape(1); b.mx = rho.shape(2);
  b.ghost = ghost_;
  for (int a = 0; a < 3; a++) b.cell_width[a] = cell_width_[a];
  auto ptr = [&](const char *name) {
    auto it = fields_.find(name);
    return it == fields_.end() ? nullptr : it->second.data();
  };
  b.density = ptr("density");
  b.velocity[0] = ptr("velocity_x");
  b.velocity[1] = ptr("velocity_y");
  b.velocity[2] = ptr("velocity_z");
  b.total_energy = ptr("total_energy");
  b.pressure = ptr("pressure");
  for (int a = 0; a < 3; a++) {
    b.bfield[a] = ptr(kBNames[a]);
    b.bfieldi[a] = ptr(kBiNames[a]);
  }
  b.internal_energy = ptr("internal_energy");
  return b;
}

// ---------------------------------------------- field groups, FluxData

Grouping::Grouping() {
  for (const char *f : {"density", "total_energy", "velocity_x", "velocity_y",
                        "velocity_z"})
    add(f, "conserved");
  for (const char *f : {"velocity_x", "velocity_y", "velocity_z", "total_energy"})
    add(f, "make_field_conservative");
}

void Grouping::add(const std::string &field, const std::string &group) {
  str_vec_t &g = groups_[group];
  if (std::find(g.begin(), g.end(), field) == g.end()) g.push_back(field);
}

void Grouping::clear(const std::string &group) { groups_.erase(group); }

bool Grouping::is_in(const std::string &field, const std::string &group) const {
  auto it = groups_.find(group);
  return it != groups_.end() &&
         std::find(it->second.begin(), it->second.end(), field) != it->second.end();
}

str_vec_t Grouping::group_list(const std::string &group) const {
  auto it = groups_.find(group);
  return it == groups_.end() ? str_vec_t() : it->second;
}

namespace cello {
Grouping *field_groups() {
  static Grouping groups;
  return &groups;
}
}  // namespace cello

void FluxData::allocate(int nx, int ny, int nz, const str_vec_t &field_list) {
  const long long n = enzoe_flux_data_size((int)field_list.size(), nx, ny, nz);
  ENZOE_ASSERT("FluxData::allocate", "invalid block size", n >= 0);
  // FaceFluxes::FaceFluxes (data_FaceFluxes.cpp:38-45)
  ENZOE_ASSERT("FluxData::allocate", "Block dimensions must be even",
               !(nx & 1) && !(ny & 1) && !(nz & 1));
  const std::array<int, 3> sz = {nx, ny, nz};
  allocated_ = true;
  if (storage_ && size_ == n && n_ == sz && fields_ == field_list) {
    // the buffer of an earlier cycle again, zeroed (the reference allocates
    // it anew every cycle)
    if (hipMemsetAsync(storage_.get(), 0, (std::size_t)n * sizeof(double),
                       (hipStream_t)g_stream) != hipSuccess)
      ENZOE_ERROR("FluxData::allocate", "hipMemsetAsync failed");
    return;
  }
  storage_ = device_alloc((std::size_t)n * sizeof(double), "FluxData::allocate");
  fields_ = field_list;
  n_ = sz;
  size_ = n;
}

void FluxData::deallocate() { allocated_ = false; }

long long FluxData::offset_(int axis, int face, int i_f, int nbr) const {
  const long long m[3] = {(long long)n_[1] * n_[2], (long long)n_[0] * n_[2],
                          (long long)n_[0] * n_[1]};
  long long o = (long long)i_f * 4 * (m[0] + m[1] + m[2]);
  for (int a = 0; a < axis; a++) o += 4 * m[a];
  return o + (long long)face * 2 * m[axis] + (nbr ? m[axis] : 0);
}

double *FluxData::block_fluxes(int axis, int face, int i_f) const {
  return flux_array() + offset_(axis, face, i_f, 0);
}

double *FluxData::neighbor_fluxes(int axis, int face, int i_f) const {
  return flux_array() + offset_(axis, face, i_f, 1);
}

// --------------------------------------------------------- reconstruction

EnzoReconstructor *EnzoReconstructor::construct_reconstructor(
    const str_vec_t &active_primitive_keys, std::string name,
    enzo_float theta_limiter) {
  // EnzoReconstructor.cpp:19-21
  ENZOE_ASSERT("EnzoReconstructor::construct_reconstructor",
               "theta_limiter must satisfy 1<=theta_limiter<=2",
               (1. <= theta_limiter) && (theta_limiter <= 2));
  const std::string formatted = lowered(name);
  if (formatted == "nn") return new EnzoReconstructorNN(active_primitive_keys);
  if (formatted == "plm" || formatted == "plm_enzo")
    return new EnzoReconstructorPLMEnzoRKLim(active_primitive_keys, theta_limiter);
  if (formatted == "plm_athena")
    return new EnzoReconstructorPLMAthenaLim(active_primitive_keys, theta_limiter);
  ENZOE_ERROR("EnzoReconstructor",
              "The only allowed solvers are NN, PLM, PLM_ENZO, & PLM_ATHENA");
}

void EnzoReconstructor::reconstruct_interface(const EnzoEFltArrayMap &prim_map,
                                              EnzoEFltArrayMap &priml_map,
                                              EnzoEFltArrayMap &primr_map,
                                              const int dim, const int stale_depth,
                                              const str_vec_t &passive_list) {
  const str_vec_t keys = concat(active_key_names_, passive_list);
  const enzoe_fmap w = select_keys(prim_map, keys).slot_map(passive_list);
  const enzoe_fmap l = select_keys(priml_map, keys).slot_map(passive_list);
  const enzoe_fmap r = select_keys(primr_map, keys).slot_map(passive_list);
  const EnzoPhysicsFluidProps *fp = enzo::fluid_props();
  check_rc(enzoe_reconstruct_interface(device_name_(), theta_(), fp->density_floor(),
                                       fp->pressure_floor(), &w, &l, &r, dim,
                                       stale_depth, g_stream),
           "EnzoReconstructor::reconstruct_interface");
}

// ---------------------------------------------------------------- Riemann

EnzoRiemann *EnzoRiemann::construct_riemann(const FactoryArgs &factory_args) noexcept {
  // EnzoRiemann.cpp:26-68
  const std::string formatted = lowered(factory_args.solver);
  if (formatted == "hll") {
    ENZOE_ASSERT("EnzoRiemann::construct_riemann",
                 "An \"HLL\" Riemann solver without magnetic fields isn't "
                 "currently supported.", factory_args.mhd);
    // the reference's Davis-wavespeed HLL aborts when used
    // (EnzoRiemannHLL.hpp:201)
    ENZOE_ERROR("EnzoRiemann::construct_riemann",
                "The \"HLL\" (Davis wavespeed) solver is not usable; use HLLE");
  } else if (formatted == "hlle") {
    if (!factory_args.mhd)
      ENZOE_ERROR("EnzoRiemann::construct_riemann",
                  "The \"HLLE\" Riemann solver without magnetic fields is untested");
    return new EnzoRiemann("hlle", true, factory_args.internal_energy);
  } else if (formatted == "hllc") {
    ENZOE_ASSERT("EnzoRiemann::construct_riemann",
                 "The \"HLLC\" Riemann Solver can't support mhd", !factory_args.mhd);
    return new EnzoRiemann("hllc", false, factory_args.internal_energy);
  } else if (formatted == "hlld") {
    ENZOE_ASSERT("EnzoRiemann::construct_riemann",
                 "The \"HLLD\" Riemann Solver requires magnetic fields",
                 factory_args.mhd);
    return new EnzoRiemann("hlld", true, factory_args.internal_energy);
  }
  ENZOE_ERROR("EnzoRiemann::construct_riemann",
              "The only known solvers are HLL, HLLE, HLLC, & HLLD");
}

EnzoRiemann::EnzoRiemann(std::string solver, bool mhd, bool internal_energy)
    : solver_(std::move(solver)), mhd_(mhd) {
  // key order of the LUTs (EnzoRiemannLUT.hpp; EnzoRiemannHLLD.hpp,
  // EnzoRiemannHLLC.hpp): density, [bfield], velocity, total_energy/pressure
  if (mhd_) {
    integration_keys_ = {"density",    "bfield_x",   "bfield_y",   "bfield_z",
                         "velocity_x", "velocity_y", "velocity_z", "total_energy"};
  } else {
    integration_keys_ = {"density", "velocity_x", "velocity_y", "velocity_z",
                         "total_energy"};
  }
  primitive_keys_ = integration_keys_;
  primitive_keys_.back() = "pressure";
  if (internal_energy) integration_keys_.push_back("internal_energy");
}

void EnzoRiemann::solve(const EnzoEFltArrayMap &prim_map_l,
                        const EnzoEFltArrayMap &prim_map_r,
                        EnzoEFltArrayMap &flux_map, const int dim,
                        const int stale_depth, const str_vec_t &passive_list,
                        const EFlt3DArray *const interface_velocity) const {
  const str_vec_t pk = concat(primitive_keys_, passive_list);
  const str_vec_t ik = concat(integration_keys_, passive_list);
  const enzoe_fmap l = select_keys(prim_map_l, pk).slot_map(passive_list);
  const enzoe_fmap r = select_keys(prim_map_r, pk).slot_map(passive_list);
  const enzoe_fmap f = select_keys(flux_map, ik).slot_map(passive_list);
  if (f.data[ENZOE_SLOT_INTERNAL_ENERGY] != nullptr) {
    // calculate_internal_energy_flux_ (EnzoRiemannImpl.hpp:278-301,
    // prep_dual_energy_arrays_ EnzoRiemannUtils.hpp:374-395)
    ENZOE_ASSERT("enzo_riemann_utils::prep_dual_energy_arrays_",
                 "interface_velocity is expected to be non-NULL when computing the "
                 "internal energy flux", interface_velocity != nullptr);
    const enzoe_farray v = interface_velocity->desc();
    check_rc(enzoe_riemann_solve_dual_energy(solver_.c_str(), mhd_ ? 1 : 0,
                                             enzo::fluid_props()->gamma(), &l, &r, &f, &v,
                                             dim, stale_depth, g_stream),
             "EnzoRiemann::solve");
    return;
  }
  check_rc(enzoe_riemann_solve(solver_.c_str(), mhd_ ? 1 : 0,
                               enzo::fluid_props()->gamma(), &l, &r, &f, dim,
                               stale_depth, g_stream),
           "EnzoRiemann::solve");
}

// ------------------------------------------------- constrained transport

EnzoBfieldMethod::EnzoBfieldMethod(int num_partial_timesteps)
    : num_partial_timesteps_(num_partial_timesteps) {
  // EnzoBfieldMethod.cpp:14-27
  ENZOE_ASSERT("EnzoConstrainedTransport", "num_partial_timesteps must be positive",
               num_partial_timesteps_ > 0);
  if (num_partial_timesteps_ != 2)
    ENZOE_ERROR("EnzoConstrainedTransport",
                "This machinery hasn't been tested for cases when "
                "num_partial_timesteps!=2.");
}

void EnzoBfieldMethod::register_target_block(Block *block) noexcept {
  // EnzoBfieldMethod.cpp:32-49
  if (target_block_ != nullptr)
    ENZOE_ERROR("EnzoConstrainedTransport::register_target_block",
                "A target block is already specified");
  target_block_ = block;
  ENZOE_ASSERT("EnzoBfieldMethod::register_target_block", "Invalid state.",
               partial_timestep_index_ <= 0);
  bool first_initialization = (partial_timestep_index_ < 0);
  register_target_block_(block, first_initialization);
  partial_timestep_index_ = 0;
}

void EnzoBfieldMethod::increment_partial_timestep() noexcept {
  // EnzoBfieldMethod.cpp:53-63
  require_registered_block_();
  if ((partial_timestep_index_ + 1) == num_partial_timesteps_) {
    partial_timestep_index_ = 0;
    target_block_ = nullptr;
  } else {
    partial_timestep_index_++;
  }
}

void EnzoBfieldMethod::require_registered_block_() const noexcept {
  ENZOE_ASSERT("EnzoBfieldMethod::require_registered_block_",
               "A target block has not been registered", target_block_ != nullptr);
}

EnzoBfieldMethodCT::EnzoBfieldMethodCT(int num_partial_timesteps)
    : EnzoBfieldMethod(num_partial_timesteps) {}

void EnzoBfieldMethodCT::register_target_block_(Block *block,
                                                bool first_initialization) noexcept {
  // EnzoBfieldMethodCT.cpp:26-77
  for (int i = 0; i < 3; i++) bfieldi_l_[i] = block->field_view(kBiNames[i]);
  cell_widths_ = block->CellWidth();
  const int mz = bfieldi_l_[0].shape(0);
  const int my = bfieldi_l_[0].shape(1);
  const int mx = bfieldi_l_[0].shape(2) - 1;
  // scratch follows the block's shape (the reference sizes it once, on the
  // first block)
  const bool reshape = center_efield_.is_null() || center_efield_.shape(0) != mz ||
                       center_efield_.shape(1) != my || center_efield_.shape(2) != mx;
  if (first_initialization || reshape) {
    if (num_partial_timesteps() > 1)
      for (int i = 0; i < 3; i++)
        temp_bfieldi_l_[i] = EFlt3DArray(bfieldi_l_[i].shape(0), bfieldi_l_[i].shape(1),
                                         bfieldi_l_[i].shape(2));
    weight_l_[0] = EFlt3DArray(mz, my, mx - 1);
    weight_l_[1] = EFlt3DArray(mz, my - 1, mx);
    weight_l_[2] = EFlt3DArray(mz - 1, my, mx);
    edge_efield_l_[0] = EFlt3DArray(mz - 1, my - 1, mx);
    edge_efield_l_[1] = EFlt3DArray(mz - 1, my, mx - 1);
    edge_efield_l_[2] = EFlt3DArray(mz, my - 1, mx - 1);
    center_efield_ = EFlt3DArray(mz, my, mx);
  }
}

void EnzoBfieldMethodCT::correct_reconstructed_bfield(EnzoEFltArrayMap &l_map,
                                                      EnzoEFltArrayMap &r_map,
                                                      int dim,
                                                      int stale_depth) noexcept {
  // EnzoBfieldMethodCT.cpp:122-166
  require_registered_block_();
  std::array<EFlt3DArray, 3> *cur_bfieldi_l;
  if (partial_timestep_index() == 0) {
    cur_bfieldi_l = &bfieldi_l_;
  } else if (partial_timestep_index() == 1) {
    cur_bfieldi_l = &temp_bfieldi_l_;
  } else {
    ENZOE_ERROR("EnzoBfieldMethodCT::correct_reconstructed_bfield",
                "Unsure how to handle current partial_timestep_index_ of " +
                    std::to_string(partial_timestep_index()));
  }
  if ((dim < 0) || (dim > 2))
    ENZOE_ERROR("EnzoBfieldMethodCT::correct_reconstructed_bfield",
                "dim has an invalid value");
  const str_vec_t key{kBNames[dim]};
  const enzoe_farray bi = (*cur_bfieldi_l)[dim].desc();
  const enzoe_fmap l = select_keys(l_map, key).slot_map();
  const enzoe_fmap r = select_keys(r_map, key).slot_map();
  check_rc(enzoe_ct_correct_reconstructed_bfield(dim, &bi, &l, &r, stale_depth, g_stream),
           "EnzoBfieldMethodCT::correct_reconstructed_bfield");
}

void EnzoBfieldMethodCT::identify_upwind(const EnzoEFltArrayMap &flux_map, int dim,
                                         int stale_depth) noexcept {
  // EnzoBfieldMethodCT.cpp:170-216
  require_registered_block_();
  if ((dim < 0) || (dim > 2))
    ENZOE_ERROR("EnzoBfieldMethodCT::identify_upwind", "dim has an invalid value");
  const enzoe_fmap f = select_keys(flux_map, {"density"}).slot_map();
  const enzoe_farray w = weight_l_[dim].desc();
  check_rc(enzoe_ct_identify_upwind(dim, &f, &w, stale_depth, g_stream),
           "EnzoBfieldMethodCT::identify_upwind");
}

void EnzoBfieldMethodCT::update_all_bfield_components(
    EnzoEFltArrayMap &cur_integration_map, const EnzoEFltArrayMap &xflux_map,
    const EnzoEFltArrayMap &yflux_map, const EnzoEFltArrayMap &zflux_map,
    EnzoEFltArrayMap &out_centered_bfield_map, enzo_float dt,
    int stale_depth) noexcept {
  // EnzoBfieldMethodCT.cpp:220-263
  require_registered_block_();
  std::array<EFlt3DArray, 3> *cur_bfieldi_l = &bfieldi_l_;
  std::array<EFlt3DArray, 3> *out_bfieldi_l;
  if ((partial_timestep_index() == 1) || (num_partial_timesteps() == 1)) {
    out_bfieldi_l = &bfieldi_l_;
  } else if (partial_timestep_index() == 0) {
    out_bfieldi_l = &temp_bfieldi_l_;
  } else {
    ENZOE_ERROR("EnzoBfieldMethodCT::correct_reconstructed_bfield",
                "Unsure how to handle current partial_timestep_index_ of " +
                    std::to_string(partial_timestep_index()));
  }
  compute_all_edge_efields(cur_integration_map, xflux_map, yflux_map, zflux_map,
                           center_efield_, edge_efield_l_, weight_l_, stale_depth);
  for (int dim = 0; dim < 3; dim++)
    update_bfield(cell_widths_, dim, edge_efield_l_, (*cur_bfieldi_l)[dim],
                  (*out_bfieldi_l)[dim], dt, stale_depth);
  for (int dim = 0; dim < 3; dim++)
    compute_center_bfield(dim, out_centered_bfield_map.at(kBNames[dim]),
                          (*out_bfieldi_l)[dim], stale_depth);
}

void EnzoBfieldMethodCT::compute_center_efield(const int dim, EFlt3DArray &efield,
                                               const EnzoEFltArrayMap &integration_map,
                                               int stale_depth) {
  // EnzoBfieldMethodCT.cpp:267-292
  const enzoe_farray e = efield.desc();
  const enzoe_fmap u =
      select_keys(integration_map, {"density", "velocity_x", "velocity_y", "velocity_z",
                                    "total_energy", "bfield_x", "bfield_y", "bfield_z"})
          .slot_map();
  check_rc(enzoe_ct_compute_center_efield(dim, &e, &u, stale_depth, g_stream),
           "EnzoBfieldMethodCT::compute_center_efield");
}

void EnzoBfieldMethodCT::compute_edge_efield(int dim, const EFlt3DArray &center_efield,
                                             const EFlt3DArray &edge_efield,
                                             const EnzoEFltArrayMap &jflux_map,
                                             const EnzoEFltArrayMap &kflux_map,
                                             const std::array<EFlt3DArray, 3> &weight_l,
                                             int stale_depth) {
  // EnzoBfieldMethodCT.cpp:480-557
  EnzoPermutedCoordinates coord(dim);
  const enzoe_farray ec = center_efield.desc(), ed = edge_efield.desc();
  const enzoe_fmap fj = select_keys(jflux_map, {kBNames[coord.k_axis()]}).slot_map();
  const enzoe_fmap fk = select_keys(kflux_map, {kBNames[coord.j_axis()]}).slot_map();
  const enzoe_farray w[3] = {weight_l[0].desc(), weight_l[1].desc(), weight_l[2].desc()};
  check_rc(enzoe_ct_compute_edge_efield(dim, &ec, &ed, &fj, &fk, w, stale_depth, g_stream),
           "EnzoBfieldMethodCT::compute_edge_efield");
}

void EnzoBfieldMethodCT::compute_all_edge_efields(
    const EnzoEFltArrayMap &integration_map, const EnzoEFltArrayMap &xflux_map,
    const EnzoEFltArrayMap &yflux_map, const EnzoEFltArrayMap &zflux_map,
    EFlt3DArray &center_efield, std::array<EFlt3DArray, 3> &edge_efield_l,
    const std::array<EFlt3DArray, 3> &weight_l, int stale_depth) {
  // EnzoBfieldMethodCT.cpp:561-590
  for (int i = 0; i < 3; i++) {
    compute_center_efield(i, center_efield, integration_map, stale_depth);
    const EnzoEFltArrayMap *jflux_map, *kflux_map;
    if (i == 0) {
      jflux_map = &yflux_map;
      kflux_map = &zflux_map;
    } else if (i == 1) {
      jflux_map = &zflux_map;
      kflux_map = &xflux_map;
    } else {
      jflux_map = &xflux_map;
      kflux_map = &yflux_map;
    }
    compute_edge_efield(i, center_efield, edge_efield_l[i], *jflux_map, *kflux_map,
                        weight_l, stale_depth);
  }
}

void EnzoBfieldMethodCT::update_bfield(const enzo_float *cell_widths, int dim,
                                       const std::array<EFlt3DArray, 3> &efield_l,
                                       const EFlt3DArray &cur_interface_bfield,
                                       const EFlt3DArray &out_interface_bfield,
                                       enzo_float dt, int stale_depth) {
  // EnzoBfieldMethodCT.cpp:617-687
  const enzoe_farray e[3] = {efield_l[0].desc(), efield_l[1].desc(), efield_l[2].desc()};
  const enzoe_farray cur = cur_interface_bfield.desc(), out = out_interface_bfield.desc();
  check_rc(enzoe_ct_update_bfield(cell_widths, dim, e, &cur, &out, dt, stale_depth,
                                  g_stream),
           "EnzoBfieldMethodCT::update_bfield");
}

void EnzoBfieldMethodCT::compute_center_bfield(int dim, const EFlt3DArray &bfieldc_comp,
                                               const EFlt3DArray &bfieldi_comp,
                                               int stale_depth) {
  // EnzoBfieldMethodCT.cpp:702-728
  const enzoe_farray c = bfieldc_comp.desc(), f = bfieldi_comp.desc();
  check_rc(enzoe_ct_compute_center_bfield(dim, &c, &f, stale_depth, g_stream),
           "EnzoBfieldMethodCT::compute_center_bfield");
}

// ---------------------------------------------------- quantity updating

EnzoIntegrationQuanUpdate::EnzoIntegrationQuanUpdate(
    const str_vec_t &integration_quantity_keys, const bool skip_B_update) throw() {
  // EnzoIntegrationQuanUpdate.cpp:61-79: conserved keys (density, and B
  // unless skipped) first, then the specific ones
  for (const std::string &k : integration_quantity_keys)
    if (k == "density") integration_keys_.push_back(k);
  if (!skip_B_update)
    for (const std::string &k : integration_quantity_keys)
      if (k.rfind("bfield_", 0) == 0) integration_keys_.push_back(k);
  for (const std::string &k : integration_quantity_keys)
    if (k.rfind("velocity_", 0) == 0 || k == "total_energy" || k == "internal_energy")
      integration_keys_.push_back(k);
  ENZOE_ASSERT("EnzoIntegrationQuanUpdate",
               "\"density\" must be a registered integration quantity.",
               !integration_keys_.empty() && integration_keys_[0] == "density");
  ENZOE_ASSERT("EnzoIntegrationQuanUpdate",
               "updating B through dUcons is not provided (CT updates B)",
               skip_B_update);
}

void EnzoIntegrationQuanUpdate::clear_dUcons_map(EnzoEFltArrayMap &dUcons_map,
                                                 enzo_float value,
                                                 const str_vec_t &passive_list) const noexcept {
  // EnzoIntegrationQuanUpdate.cpp:83-101
  const enzoe_fmap m =
      select_keys(dUcons_map, concat(integration_keys_, passive_list)).slot_map(passive_list);
  check_rc(enzoe_fill_map(&m, value, g_stream), "EnzoIntegrationQuanUpdate::clear_dUcons_map");
}

void EnzoIntegrationQuanUpdate::accumulate_flux_component(
    int dim, double dt, enzo_float cell_width, const EnzoEFltArrayMap &flux_map,
    EnzoEFltArrayMap &dUcons_map, int stale_depth,
    const str_vec_t &passive_list) const noexcept {
  // EnzoIntegrationQuanUpdate.cpp:105-147
  const str_vec_t keys = concat(integration_keys_, passive_list);
  const enzoe_fmap f = select_keys(flux_map, keys).slot_map(passive_list);
  const enzoe_fmap d = select_keys(dUcons_map, keys).slot_map(passive_list);
  check_rc(enzoe_accumulate_flux_component(dim, dt, cell_width, &f, &d, stale_depth, g_stream),
           "EnzoIntegrationQuanUpdate::accumulate_flux_component");
}

void EnzoIntegrationQuanUpdate::update_quantities(EnzoEFltArrayMap &initial_integration_map,
                                                  const EnzoEFltArrayMap &dUcons_map,
                                                  EnzoEFltArrayMap &out_integration_map,
                                                  const int stale_depth,
                                                  const str_vec_t &passive_list) const {
  // EnzoIntegrationQuanUpdate.cpp:183-238 + apply_floor_to_energy_and_sync
  const str_vec_t keys = concat(integration_keys_, passive_list);
  const enzoe_fmap a = select_keys(initial_integration_map, keys).slot_map(passive_list);
  const enzoe_fmap d = select_keys(dUcons_map, keys).slot_map(passive_list);
  const bool mhd = out_integration_map.contains("bfield_x");
  str_vec_t out_keys = keys;
  if (mhd) out_keys.insert(out_keys.end(), {"bfield_x", "bfield_y", "bfield_z"});
  const enzoe_fmap o = select_keys(out_integration_map, out_keys).slot_map(passive_list);
  const EnzoPhysicsFluidProps *fp = enzo::fluid_props();
  if (fp->dual_energy_enabled())
    // internal_energy is among the specific keys; the synchronisation
    // replaces the energy floor
    check_rc(enzoe_update_quantities_dual_energy(&a, &d, &o, mhd ? 1 : 0, fp->gamma(),
                                                 fp->density_floor(), fp->pressure_floor(),
                                                 fp->dual_energy_eta(), stale_depth,
                                                 g_stream),
             "EnzoIntegrationQuanUpdate::update_quantities");
  else
    check_rc(enzoe_update_quantities(&a, &d, &o, mhd ? 1 : 0, fp->gamma(),
                                     fp->density_floor(), fp->pressure_floor(), stale_depth,
                                     g_stream),
             "EnzoIntegrationQuanUpdate::update_quantities");
}

// ------------------------------------------------------- stage commands

EnzoMHDIntegratorStageCommands::EnzoMHDIntegratorStageCommands(
    const EnzoMHDIntegratorStageArgPack &args)
    : mhd_choice_(parse_bfield_choice_(args.mhd_choice)) {
  // EnzoMHDIntegratorStageCommands.cpp:18-69
  const EnzoPhysicsFluidProps *fp = enzo::fluid_props();
  ENZOE_ASSERT("EnzoMethodMHDVlct::EnzoMethodMHDVlct",
               "selected formulation of dual energy formalism is incompatible",
               !fp->dual_energy_enabled() || fp->dual_energy_modern());
  riemann_solver_ = EnzoRiemann::construct_riemann(
      {args.rsolver, mhd_choice_ != bfield_choice::no_bfield,
       enzo::fluid_props()->dual_energy_enabled()});
  const str_vec_t primitive_field_list = riemann_solver_->primitive_quantity_keys();
  for (const std::string &name : args.recon_names)
    reconstructors_.emplace_back(EnzoReconstructor::construct_reconstructor(
        primitive_field_list, name, (enzo_float)args.theta_limiter));
  integration_quan_updater_ =
      new EnzoIntegrationQuanUpdate(riemann_solver_->integration_quantity_keys(), true);
}

EnzoMHDIntegratorStageCommands::~EnzoMHDIntegratorStageCommands() {
  delete riemann_solver_;
  delete integration_quan_updater_;
}

EnzoMHDIntegratorStageCommands::bfield_choice
EnzoMHDIntegratorStageCommands::parse_bfield_choice_(std::string choice) noexcept {
  // EnzoMHDIntegratorStageCommands.cpp:76-98
  const std::string formatted = lowered(choice);
  if (formatted == "no_bfield") return bfield_choice::no_bfield;
  if (formatted == "unsafe_constant_uniform")
    ENZOE_ERROR("EnzoMHDIntegratorStageCommands::parse_bfield_choice_",
                "constant_uniform is primarilly for debugging purposes. DON'T use "
                "for science runs (things can break).");
  if (formatted == "constrained_transport") return bfield_choice::constrained_transport;
  ENZOE_ERROR("EnzoMHDIntegratorStageCommands::parse_bfield_choice_",
              "Unrecognized choice. Known options include \"no_bfield\" and "
              "\"constrained_transport\"");
}

void EnzoMHDIntegratorStageCommands::check_valid_stage_index_(int stage_index) const noexcept {
  ENZOE_ASSERT("EnzoMHDIntegratorStageCommands::check_valid_stage_index_",
               "invalid stage index",
               stage_index >= 0 && stage_index < (int)reconstructors_.size());
}

int EnzoMHDIntegratorStageCommands::staling_from_stage(int stage_index) const noexcept {
  // EnzoMHDIntegratorStageCommands.cpp:242-245
  check_valid_stage_index_(stage_index);
  return reconstructors_[(std::size_t)stage_index]->total_staling_rate();
}

EnzoBfieldMethod *EnzoMHDIntegratorStageCommands::construct_bfield_method(
    int num_partial_timesteps) const noexcept {
  // EnzoMHDIntegratorStageCommands.hpp:116-122
  if (mhd_choice_ == bfield_choice::constrained_transport)
    return new EnzoBfieldMethodCT(num_partial_timesteps);
  return nullptr;
}

str_vec_t EnzoMHDIntegratorStageCommands::integration_quantity_keys() const noexcept {
  return riemann_solver_->integration_quantity_keys();
}

str_vec_t EnzoMHDIntegratorStageCommands::primitive_quantity_keys() const noexcept {
  return riemann_solver_->primitive_quantity_keys();
}

str_vec_t EnzoMHDIntegratorStageCommands::dUcons_map_keys() const noexcept {
  return integration_quan_updater_->integration_keys();
}

void EnzoMHDIntegratorStageCommands::compute_update_stage(
    EnzoEFltArrayMap tstep_begin_integration_map,
    EnzoEFltArrayMap cur_stage_integration_map, EnzoEFltArrayMap out_integration_map,
    EnzoEFltArrayMap primitive_map, EnzoEFltArrayMap priml_map,
    EnzoEFltArrayMap primr_map, std::array<EnzoEFltArrayMap, 3> flux_maps_xyz,
    EnzoEFltArrayMap dUcons_map, const EnzoEFltArrayMap &accel_map,
    EFlt3DArray interface_vel_arr, const str_vec_t &passive_list,
    EnzoBfieldMethod *bfield_method, unsigned short stage_index, double cur_dt,
    int stale_depth, const std::array<enzo_float, 3> cell_widths_xyz) const noexcept {
  // EnzoMHDIntegratorStageCommands.cpp:102-203
  check_valid_stage_index_((int)stage_index);
  EnzoReconstructor *reconstructor = reconstructors_[stage_index].get();
  integration_quan_updater_->clear_dUcons_map(dUcons_map, 0., passive_list);
  enzo::fluid_props()->primitive_from_integration(cur_stage_integration_map, primitive_map,
                                                  stale_depth, passive_list, true);
  for (int dim = 0; dim < 3; dim++) {
    // trimmed views of the cell-shaped reconstruction scratch
    CSlice x_slc = (dim == 0) ? CSlice(0, -1) : CSlice(0, nullptr);
    CSlice y_slc = (dim == 1) ? CSlice(0, -1) : CSlice(0, nullptr);
    CSlice z_slc = (dim == 2) ? CSlice(0, -1) : CSlice(0, nullptr);
    EnzoEFltArrayMap pl_map = priml_map.subarray_map(z_slc, y_slc, x_slc);
    EnzoEFltArrayMap pr_map = primr_map.subarray_map(z_slc, y_slc, x_slc);
    // dual energy: the interface velocities of this dim, trimmed like the
    // reconstructed states (:156-169); no scratch otherwise
    EFlt3DArray sliced_interface_vel_arr;
    const EFlt3DArray *interface_vel_arr_ptr = nullptr;
    if (enzo::fluid_props()->dual_energy_enabled()) {
      sliced_interface_vel_arr = interface_vel_arr.subarray(z_slc, y_slc, x_slc);
      interface_vel_arr_ptr = &sliced_interface_vel_arr;
    }
    compute_flux_(dim, cur_dt, cell_widths_xyz[dim], primitive_map, pl_map, pr_map,
                  flux_maps_xyz[dim], dUcons_map, interface_vel_arr_ptr, *reconstructor,
                  bfield_method, stale_depth, passive_list);
  }
  stale_depth += reconstructor->immediate_staling_rate();
  // compute_source_terms_ (EnzoMHDIntegratorStageCommands.cpp:265-295):
  // gravity on the stage with index 1, from the start-of-step state
  if ((stage_index == 1) && (accel_map.size() != 0)) {
    EnzoSourceGravity gravity_source;
    gravity_source.calculate_source(cur_dt, tstep_begin_integration_map, dUcons_map,
                                    accel_map, stale_depth);
  }
  if (bfield_method != nullptr)
    bfield_method->update_all_bfield_components(cur_stage_integration_map,
                                                flux_maps_xyz[0], flux_maps_xyz[1],
                                                flux_maps_xyz[2], out_integration_map,
                                                cur_dt, stale_depth);
  integration_quan_updater_->update_quantities(tstep_begin_integration_map, dUcons_map,
                                               out_integration_map, stale_depth,
                                               passive_list);
}

void EnzoSourceInternalEnergy::calculate_source(const int dim, const double dt,
                                                const enzo_float cell_width,
                                                const EnzoEFltArrayMap &prim_map,
                                                EnzoEFltArrayMap &dUcons_map,
                                                const EFlt3DArray &interface_velocity,
                                                const int stale_depth) const throw() {
  // EnzoSourceInternalEnergy.cpp:16-92
  const EnzoPhysicsFluidProps *fp = enzo::fluid_props();
  ENZOE_ASSERT("EnzoSourceInternalEnergy::calculate_source",
               "This function should only be called if the \"modern formulation\" of the "
               "dual energy formalism is in use.",
               fp->dual_energy_modern());
  const enzoe_fmap w = select_keys(prim_map, {"pressure"}).slot_map();
  const enzoe_fmap d = select_keys(dUcons_map, {"internal_energy"}).slot_map();
  const enzoe_farray v = interface_velocity.desc();
  check_rc(enzoe_source_internal_energy(dim, dt, cell_width, &w, &d, &v,
                                        fp->pressure_floor(), stale_depth, g_stream),
           "EnzoSourceInternalEnergy::calculate_source");
}

void EnzoSourceGravity::calculate_source(const double dt, const EnzoEFltArrayMap &prim_map,
                                         EnzoEFltArrayMap &dUcons_map,
                                         const EnzoEFltArrayMap &accel_map,
                                         const int stale_depth) const noexcept {
  const enzoe_fmap u = select_keys(prim_map, {"density", "velocity_x", "velocity_y",
                                              "velocity_z"}).slot_map();
  const enzoe_fmap d = select_keys(dUcons_map, {"velocity_x", "velocity_y", "velocity_z",
                                                "total_energy"}).slot_map();
  // acceleration_x..z in the velocity slots
  enzoe_fmap a;
  std::memset(&a, 0, sizeof(a));
  const char *names[3] = {"acceleration_x", "acceleration_y", "acceleration_z"};
  for (int i = 0; i < 3; i++) {
    const EFlt3DArray arr = accel_map.at(names[i]);
    a.data[ENZOE_SLOT_VELOCITY_X + i] = arr.data();
    for (int k = 0; k < 3; k++) {
      if (i == 0) {
        a.shape[k] = arr.shape(k);
        a.stride[k] = arr.stride(k);
      }
    }
  }
  check_rc(enzoe_source_gravity(dt, &u, &d, &a, stale_depth, g_stream),
           "EnzoSourceGravity::calculate_source");
}

void EnzoMHDIntegratorStageCommands::compute_flux_(
    const int dim, const double cur_dt, const enzo_float cell_width,
    EnzoEFltArrayMap &primitive_map, EnzoEFltArrayMap &priml_map,
    EnzoEFltArrayMap &primr_map, EnzoEFltArrayMap &flux_map,
    EnzoEFltArrayMap &dUcons_map, const EFlt3DArray *const interface_velocity_arr_ptr,
    EnzoReconstructor &reconstructor, EnzoBfieldMethod *bfield_method,
    const int stale_depth, const str_vec_t &passive_list) const noexcept {
  // EnzoMHDIntegratorStageCommands.cpp:207-261
  reconstructor.reconstruct_interface(primitive_map, priml_map, primr_map, dim,
                                      stale_depth, passive_list);
  int cur_stale_depth = stale_depth + reconstructor.immediate_staling_rate();
  if (bfield_method != nullptr)
    bfield_method->correct_reconstructed_bfield(priml_map, primr_map, dim, cur_stale_depth);
  riemann_solver_->solve(priml_map, primr_map, flux_map, dim, cur_stale_depth,
                         passive_list, interface_velocity_arr_ptr);
  integration_quan_updater_->accumulate_flux_component(dim, cur_dt, cell_width, flux_map,
                                                       dUcons_map, cur_stale_depth,
                                                       passive_list);
  // the internal-energy source term of this dim (:248-255)
  if (enzo::fluid_props()->dual_energy_enabled()) {
    EnzoSourceInternalEnergy eint_src;
    eint_src.calculate_source(dim, cur_dt, cell_width, primitive_map, dUcons_map,
                              *interface_velocity_arr_ptr, cur_stale_depth);
  }
  if (bfield_method != nullptr)
    bfield_method->identify_upwind(flux_map, dim, cur_stale_depth);
}

double EnzoMHDIntegratorStageCommands::timestep(EnzoEFltArrayMap integration_map,
                                                const EFlt3DArray &pressure,
                                                const double dx, const double dy,
                                                const double dz) const noexcept {
  // EnzoMHDIntegratorStageCommands.cpp:299-366
  str_vec_t keys = {"density", "velocity_x", "velocity_y", "velocity_z", "total_energy"};
  if (!is_pure_hydro()) keys.insert(keys.end(), {"bfield_x", "bfield_y", "bfield_z"});
  const enzoe_fmap u = select_keys(integration_map, keys).slot_map();
  const enzoe_farray p = pressure.desc();
  double dt = 0.;
  check_rc(enzoe_integrator_timestep(&u, &p, is_pure_hydro() ? 0 : 1,
                                     enzo::fluid_props()->gamma(), dx, dy, dz, g_stream, &dt),
           "EnzoMHDIntegratorStageCommands::timestep");
  return dt;
}

EnzoVlctScratchSpace::EnzoVlctScratchSpace(const std::array<int, 3> &shape,
                                           const str_vec_t &integration_keys,
                                           const str_vec_t &primitive_keys,
                                           const str_vec_t &dUcons_keys,
                                           const str_vec_t &passives, bool dual_energy)
    : temp_integration_map("temp_integration", concat(integration_keys, passives), shape),
      primitive_map("primitive", concat(primitive_keys, passives), shape),
      priml_map("priml", concat(primitive_keys, passives), shape),
      primr_map("primr", concat(primitive_keys, passives), shape),
      xflux_map("xflux", concat(integration_keys, passives),
                std::array<int, 3>{{shape[0], shape[1], shape[2] - 1}}),
      yflux_map("yflux", concat(integration_keys, passives),
                std::array<int, 3>{{shape[0], shape[1] - 1, shape[2]}}),
      zflux_map("zflux", concat(integration_keys, passives),
                std::array<int, 3>{{shape[0] - 1, shape[1], shape[2]}}),
      dUcons_map("dUcons", concat(dUcons_keys, passives), shape),
      passive_list(passives) {
  if (dual_energy) interface_vel_arr = EFlt3DArray(shape[0], shape[1], shape[2]);
}

// ----------------------------------------------------------------- method

EnzoMethodMHDVlct::EnzoMethodMHDVlct(ParameterGroup p, bool store_fluxes_for_corrections) {
  // parse_mhdchoice_pack_pair_ (EnzoMethodMHDVlct.cpp:27-86)
  auto is_defined = [&](const std::string &n) { return p.param(n) != nullptr; };
  time_scheme_ = p.value_string("time_scheme", "vl");
  str_vec_t recon_names;
  if (time_scheme_ == "euler") {
    recon_names = {p.value_string("reconstruct_method", "plm")};
    std::fprintf(stderr,
                 "WARNING in parse_mhdchoice_pack_pair_: \"euler\" temporal "
                 "integration exists for debugging purposes. It has not been "
                 "rigorously tested. USE WITH CAUTION\n");
  } else if (time_scheme_ == "vl") {
    recon_names = {"nn", p.value_string("reconstruct_method", "plm")};
  } else {
    ENZOE_ERROR("parse_mhdchoice_pack_pair_",
                "\"" + p.get_group_path() + ":time_scheme\" must be \"vl\" or \"euler\"");
  }
  if (is_defined("half_dt_reconstruct_method") || is_defined("full_dt_reconstruct_method"))
    ENZOE_ERROR("parse_mhdchoice_pack_pair_",
                "In the \"" + p.get_group_path() +
                    "\" parameter-group, \"half_dt_reconstruct_method\" & "
                    "\"full_dt_reconstruct_method\" have been removed. The former was "
                    "only allowed to have a value of \"nn\" and the latter was "
                    "replaced with \"reconstruct_method\".");
  if (!is_defined("mhd_choice"))
    ENZOE_ERROR("parse_mhdchoice_pack_pair_", p.full_name("mhd_choice") + " wasn't specified");
  integrator_arg_pack_ = new EnzoMHDIntegratorStageArgPack{
      p.value_string("riemann_solver", "hlld"), recon_names,
      p.value_float("theta_limiter", 1.5), p.value_string("mhd_choice", "")};

  // EnzoMethodMHDVlct.cpp:94-150
  const double dflt_courant = (time_scheme_ == "vl") ? 0.3 : 1.0;
  set_courant(p.value_float("courant", dflt_courant));
  const int nstages = (int)integrator_arg_pack_->recon_names.size();
  integrator_ = new EnzoMHDIntegratorStageCommands(*integrator_arg_pack_);
  integration_field_list_ = integrator_->integration_quantity_keys();
  primitive_field_list_ = integrator_->primitive_quantity_keys();
  bfield_method_ = integrator_->construct_bfield_method(nstages);
  if (bfield_method_ != nullptr) bfield_method_->check_required_fields();
  // EnzoMethodMHDVlct.cpp:136-141
  store_fluxes_for_corrections_ = store_fluxes_for_corrections;
  if (store_fluxes_for_corrections) {
    ENZOE_ASSERT("EnzoMethodMHDVlct::EnzoMethodMHDVlct",
                 "Flux corrections are currently only supported in hydro-mode",
                 integrator_->is_pure_hydro());
  }

  // device extensions
  const std::string fused = lowered(p.value_string("device_fused", "true"));
  ENZOE_ASSERT("EnzoMethodMHDVlct", p.full_name("device_fused") + " must be true or false",
               fused == "true" || fused == "false");
  fused_ = (fused == "true");
  const std::string dbg = lowered(p.value_string("debug_checks", "false"));
  ENZOE_ASSERT("EnzoMethodMHDVlct", p.full_name("debug_checks") + " must be true or false",
               dbg == "true" || dbg == "false");
  debug_checks_ = (dbg == "true");
  const std::string prec = lowered(p.value_string("precision", "double"));
  if (prec == "double") {
    precision_ = ENZOE_PREC_DOUBLE;
  } else if (prec == "mixed") {
    ENZOE_ASSERT("EnzoMethodMHDVlct",
                 "mixed precision is only provided by the fused device path", fused_);
    precision_ = ENZOE_PREC_MIXED;
  } else if (prec == "double_fma") {
    ENZOE_ASSERT("EnzoMethodMHDVlct",
                 "contracted FP64 is only provided by the fused device path", fused_);
    precision_ = ENZOE_PREC_DOUBLE_FMA;
  } else {
    ENZOE_ERROR("EnzoMethodMHDVlct", p.full_name("precision") +
                                         " must be \"double\", \"double_fma\" or \"mixed\"");
  }

  // the fused library handle (validation only; device memory on first use)
  const EnzoPhysicsFluidProps *fp = enzo::fluid_props();
  const std::string recon = recon_names.back();
  enzoe_vlct_params prm;
  std::memset(&prm, 0, sizeof(prm));
  prm.mhd_choice = integrator_arg_pack_->mhd_choice.c_str();
  prm.riemann_solver = integrator_arg_pack_->rsolver.c_str();
  prm.reconstruct_method = recon.c_str();
  prm.time_scheme = time_scheme_.c_str();
  prm.theta_limiter = integrator_arg_pack_->theta_limiter;
  prm.courant = courant_;
  prm.gamma = fp->gamma();
  prm.density_floor = fp->density_floor();
  prm.pressure_floor = fp->pressure_floor();
  prm.precision = precision_;
  if (fp->dual_energy_enabled()) {
    // the FP32 flux path has no internal-energy terms
    ENZOE_ASSERT("EnzoMethodMHDVlct",
                 "the dual-energy formalism needs " + p.full_name("precision") +
                     " = \"double\" or \"double_fma\"", precision_ != ENZOE_PREC_MIXED);
    prm.dual_energy = "modern";
    prm.dual_energy_eta = fp->dual_energy_eta();
  } else {
    prm.dual_energy = "disabled";
    prm.dual_energy_eta = -1.0;
  }
  check_rc(enzoe_vlct_create(&prm, &device_), "EnzoMethodMHDVlct::EnzoMethodMHDVlct");
}

EnzoMethodMHDVlct::~EnzoMethodMHDVlct() {
  if (device_) enzoe_vlct_destroy(device_);
  delete integrator_arg_pack_;
  delete integrator_;
  delete scratch_space_;
  delete bfield_method_;
}

EnzoEFltArrayMap EnzoMethodMHDVlct::get_integration_map_(
    Block *block, const str_vec_t &passive_list) const noexcept {
  // EnzoMethodMHDVlct.cpp:201-216 (passive scalars as densities)
  const str_vec_t field_list = concat(integration_field_list_, passive_list);
  std::vector<EFlt3DArray> arrays;
  arrays.reserve(field_list.size());
  for (const std::string &name : field_list) arrays.push_back(block->field_view(name));
  return EnzoEFltArrayMap("integration", field_list, arrays);
}

EnzoEFltArrayMap EnzoMethodMHDVlct::get_accel_map_(Block *block) noexcept {
  if (!block->has_field("acceleration_x"))
    return EnzoEFltArrayMap("accel", str_vec_t(), std::vector<EFlt3DArray>());
  const str_vec_t keys = {"acceleration_x", "acceleration_y", "acceleration_z"};
  std::vector<EFlt3DArray> arrays;
  for (const std::string &k : keys) arrays.push_back(block->field_view(k));
  return EnzoEFltArrayMap("accel", keys, arrays);
}

EnzoVlctScratchSpace *EnzoMethodMHDVlct::get_scratch_ptr_(
    const std::array<int, 3> &shape, const str_vec_t &passive_list) noexcept {
  // EnzoMethodMHDVlct.cpp:236-246 (re-made if a block of another shape or
  // with other passive scalars comes)
  if (scratch_space_ != nullptr &&
      (scratch_space_->primitive_map.array_shape(0) != shape[0] ||
       scratch_space_->primitive_map.array_shape(1) != shape[1] ||
       scratch_space_->primitive_map.array_shape(2) != shape[2] ||
       scratch_space_->passive_list != passive_list)) {
    delete scratch_space_;
    scratch_space_ = nullptr;
  }
  if (scratch_space_ == nullptr)
    scratch_space_ = new EnzoVlctScratchSpace(shape, integration_field_list_,
                                              primitive_field_list_,
                                              integrator_->dUcons_map_keys(),
                                              passive_list,
                                              enzo::fluid_props()->dual_energy_enabled());
  return scratch_space_;
}

// The debug counters of the fused step since the last check: a non-finite
// value ends the run as CelloView's CHECK_IF_FINITE (view_CelloView.hpp:
// 209-215), an applied floor as enzo_utils::apply_floor under
// RAISE_FLOOR_ERROR (utils.hpp:105-108).
void EnzoMethodMHDVlct::debug_check_(const char *where) {
  unsigned long long c[ENZOE_DEBUG_COUNTS];
  check_rc(enzoe_vlct_debug_counts(device_, g_stream, c), where);
  if (c[ENZOE_DEBUG_NONFINITE])
    ENZOE_ERROR("CelloView", "Non-Finite Value (" + std::to_string(c[ENZOE_DEBUG_NONFINITE]) +
                                 " updated cell values)");
  const unsigned long long floors = c[ENZOE_DEBUG_DENSITY_FLOOR] + c[ENZOE_DEBUG_ENERGY_FLOOR] +
                                    c[ENZOE_DEBUG_RECONSTRUCTION_FLOOR];
  if (floors)
    ENZOE_ERROR("enzo_utils::apply_floor",
                "DEBUG-MODE ERROR: applying a floor. (density " +
                    std::to_string(c[ENZOE_DEBUG_DENSITY_FLOOR]) + ", energy " +
                    std::to_string(c[ENZOE_DEBUG_ENERGY_FLOOR]) + ", reconstruction " +
                    std::to_string(c[ENZOE_DEBUG_RECONSTRUCTION_FLOOR]) + ")");
}

void EnzoMethodMHDVlct::compute(Block *block) throw() {
  if (debug_checks_ && fused_ && !debug_on_) {
    check_rc(enzoe_vlct_set_debug_checks(device_, 1), "EnzoMethodMHDVlct::compute");
    debug_on_ = true;
  }
  if (block->is_leaf()) {
    // EnzoMethodMHDVlct.cpp:124-133: ghost depth covers the staling
    int min_gdepth_req = 0;
    for (int s = 0; s < (int)integrator_arg_pack_->recon_names.size(); s++)
      min_gdepth_req += integrator_->staling_from_stage(s);
    ENZOE_ASSERT("EnzoMethodMHDVlct::compute",
                 "ghost depth must be at least " + std::to_string(min_gdepth_req) + ".",
                 block->ghost_depth() >= min_gdepth_req);
    if (store_fluxes_for_corrections_) allocate_FC_flux_buffer_(block);
    if (fused_) {
      ENZOE_ASSERT("EnzoMethodMHDVlct::compute",
                   "passively advected scalars need device_fused = false",
                   block->passive_scalars().empty());
      const enzoe_block b = block->desc();
      if (store_fluxes_for_corrections_) {
        const std::vector<int> slots = flux_slots_(*block->flux_data());
        double *fd = block->flux_data()->flux_array();
        check_rc(enzoe_vlct_compute_store_fluxes(device_, &b, 1, block->dt(),
                                                 (int)slots.size(), slots.data(), &fd,
                                                 g_stream),
                 "EnzoMethodMHDVlct::compute");
      } else if (block->has_field("acceleration_x")) {
        double *acc[3] = {block->field_view("acceleration_x").data(),
                          block->field_view("acceleration_y").data(),
                          block->field_view("acceleration_z").data()};
        check_rc(enzoe_vlct_compute_gravity(device_, &b, 1, block->dt(), acc, g_stream),
                 "EnzoMethodMHDVlct::compute");
      } else {
        check_rc(enzoe_vlct_compute(device_, &b, 1, block->dt(), g_stream),
                 "EnzoMethodMHDVlct::compute");
      }
      if (debug_on_) debug_check_("EnzoMethodMHDVlct::compute");
    } else {
      compute_staged_(block);
    }
  }
  block->compute_done();
}

void EnzoMethodMHDVlct::allocate_FC_flux_buffer_(Block *block) const {
  // EnzoMethodMHDVlct.cpp:333-355: the "conserved" group, every cycle
  const std::array<int, 3> n = block->active_size();
  block->flux_data()->allocate(n[0], n[1], n[2],
                               cello::field_groups()->group_list("conserved"));
}

std::vector<int> EnzoMethodMHDVlct::flux_slots_(const FluxData &fd) const {
  std::vector<int> slots;
  for (const std::string &name : fd.field_list()) {
    const int slot = key_slot(name);
    ENZOE_ASSERT("EnzoMethodMHDVlct::save_fluxes_for_corrections_",
                 "the fused step saves fluxes of density, velocity and total_energy only ("
                 "\"" + name + "\" is in the \"conserved\" group)",
                 slot >= ENZOE_SLOT_DENSITY && slot <= ENZOE_SLOT_ENERGY);
    slots.push_back(slot);
  }
  return slots;
}

void EnzoMethodMHDVlct::compute_blocks(const std::vector<Block *> &blocks) throw() {
  if (blocks.empty()) return;
  if (!fused_) {
    for (Block *b : blocks) compute(b);
    return;
  }
  if (debug_checks_ && !debug_on_) {
    check_rc(enzoe_vlct_set_debug_checks(device_, 1), "EnzoMethodMHDVlct::compute_blocks");
    debug_on_ = true;
  }
  std::vector<enzoe_block> desc;
  std::vector<double *> fds;
  desc.reserve(blocks.size());
  for (Block *b : blocks) {
    ENZOE_ASSERT("EnzoMethodMHDVlct::compute_blocks",
                 "blocks of one batched launch share dt", b->dt() == blocks[0]->dt());
    ENZOE_ASSERT("EnzoMethodMHDVlct::compute_blocks",
                 "passively advected scalars need device_fused = false",
                 b->passive_scalars().empty());
    desc.push_back(b->desc());
    if (store_fluxes_for_corrections_) {
      allocate_FC_flux_buffer_(b);
      fds.push_back(b->flux_data()->flux_array());
    }
  }
  if (store_fluxes_for_corrections_) {
    const std::vector<int> slots = flux_slots_(*blocks[0]->flux_data());
    check_rc(enzoe_vlct_compute_store_fluxes(device_, desc.data(), (int)desc.size(),
                                             blocks[0]->dt(), (int)slots.size(),
                                             slots.data(), fds.data(), g_stream),
             "EnzoMethodMHDVlct::compute_blocks");
  } else {
    check_rc(enzoe_vlct_compute(device_, desc.data(), (int)desc.size(), blocks[0]->dt(),
                                g_stream),
             "EnzoMethodMHDVlct::compute_blocks");
  }
  if (debug_on_) debug_check_("EnzoMethodMHDVlct::compute_blocks");
  for (Block *b : blocks) b->compute_done();
}

void EnzoMethodMHDVlct::compute_staged_(Block *block) noexcept {
  // EnzoMethodMHDVlct.cpp:371-497
  const str_vec_t passive_list = block->passive_scalars();
  EnzoEFltArrayMap external_integration_map = get_integration_map_(block, passive_list);
  const std::array<int, 3> shape = {external_integration_map.array_shape(0),
                                    external_integration_map.array_shape(1),
                                    external_integration_map.array_shape(2)};
  EnzoVlctScratchSpace *const scratch = get_scratch_ptr_(shape, passive_list);
  EnzoEFltArrayMap temp_integration_map = scratch->temp_integration_map;
  EnzoEFltArrayMap primitive_map = scratch->primitive_map;
  EnzoEFltArrayMap priml_map = scratch->priml_map;
  EnzoEFltArrayMap primr_map = scratch->primr_map;
  std::array<EnzoEFltArrayMap, 3> flux_maps_xyz = {scratch->xflux_map, scratch->yflux_map,
                                                   scratch->zflux_map};
  EnzoEFltArrayMap dUcons_map = scratch->dUcons_map;
  const EnzoEFltArrayMap accel_map = get_accel_map_(block);
  if (bfield_method_ != nullptr) bfield_method_->register_target_block(block);
  const std::array<enzo_float, 3> cell_widths_xyz = {
      block->CellWidth()[0], block->CellWidth()[1], block->CellWidth()[2]};
  const double dt = block->dt();
  int stale_depth = 0;
  unsigned short nstages;
  if (time_scheme_ == "euler") {
    nstages = 1;
  } else if (time_scheme_ == "vl") {
    nstages = 2;
  } else {
    ENZOE_ERROR("EnzoMethodMHDVlct::compute", "unrecognized time_scheme_");
  }
  for (unsigned short stage_index = 0; stage_index < nstages; stage_index++) {
    const bool is_final_stage = (stage_index + 1) == nstages;
    const double cur_dt = (!is_final_stage) ? dt / 2. : dt;
    EnzoEFltArrayMap cur_stage_integration_map =
        (stage_index == 0) ? external_integration_map : temp_integration_map;
    EnzoEFltArrayMap out_integration_map =
        (is_final_stage) ? external_integration_map : temp_integration_map;
    integrator_->compute_update_stage(external_integration_map, cur_stage_integration_map,
                                      out_integration_map, primitive_map, priml_map,
                                      primr_map, flux_maps_xyz, dUcons_map, accel_map,
                                      scratch->interface_vel_arr, passive_list,
                                      bfield_method_, stage_index, cur_dt, stale_depth,
                                      cell_widths_xyz);
    stale_depth += integrator_->staling_from_stage((int)stage_index);
    if (is_final_stage && store_fluxes_for_corrections_) {
      // EnzoMethodMHDVlct.cpp:480-489
      FluxData *fd = block->flux_data();
      std::vector<int> slots;
      for (const std::string &name : fd->field_list()) {
        int slot = key_slot(name);
        auto it = std::find(passive_list.begin(), passive_list.end(), name);
        if (slot < 0 && it != passive_list.end())
          slot = ENZOE_SLOT_PASSIVE0 + (int)(it - passive_list.begin());
        ENZOE_ASSERT("EnzoMethodMHDVlct::save_fluxes_for_corrections_",
                     "\"" + name + "\" of the \"conserved\" group has no flux", slot >= 0);
        slots.push_back(slot);
      }
      for (int i = 0; i < 3; i++) {
        const enzoe_fmap f = flux_maps_xyz[i].slot_map(passive_list);
        check_rc(enzoe_save_fluxes_for_corrections(&f, i, cell_widths_xyz[i], cur_dt,
                                                   block->ghost_depth(), (int)slots.size(),
                                                   slots.data(), fd->flux_array(),
                                                   g_stream),
                 "EnzoMethodMHDVlct::save_fluxes_for_corrections_");
      }
    }
    if (bfield_method_ != nullptr) bfield_method_->increment_partial_timestep();
  }
}

double EnzoMethodMHDVlct::timestep(Block *block) throw() {
  // EnzoMethodMHDVlct.cpp:551-588
  if (fused_) {
    const enzoe_block b = block->desc();
    double dt = 0.;
    check_rc(enzoe_vlct_timestep(device_, &b, 1, g_stream, &dt),
             "EnzoMethodMHDVlct::timestep");
    return dt;  // the library applies the same courant factor
  }
  EnzoEFltArrayMap integration_map = get_integration_map_(block, str_vec_t());
  EnzoPhysicsFluidProps *fluid_props = enzo::fluid_props();
  // synchronise eint and etot (EnzoMethodMHDVlct.cpp:563-568)
  if (fluid_props->dual_energy_enabled())
    fluid_props->apply_floor_to_energy_and_sync(integration_map, 0);
  EFlt3DArray pressure = block->field_view("pressure");
  fluid_props->pressure_from_integration(integration_map, pressure, 0);
  const double *w = block->CellWidth();
  double dtBaryons = integrator_->timestep(integration_map, pressure, w[0], w[1], w[2]);
  return dtBaryons * courant_;
}

// ------------------------------------------------ coarse/fine transfers

void ProlongLinear::apply(precision_type precision, void *values_f, int m3_f[3],
                          int o3_f[3], int n3_f[3], const void *values_c, int m3_c[3],
                          int o3_c[3], int n3_c[3], bool accumulate) {
  ENZOE_ASSERT("ProlongLinear::apply()",
               "Unknown precision " + std::to_string(precision) +
                   " (double precision only)",
               precision == precision_double || precision == precision_default);
  check_rc(enzoe_prolong_linear(static_cast<double *>(values_f), m3_f, o3_f, n3_f,
                                static_cast<const double *>(values_c), m3_c, o3_c, n3_c,
                                accumulate ? 1 : 0, g_stream),
           "ProlongLinear::apply_");
}

EnzoProlong::EnzoProlong(std::string type, bool positive, bool use_linear)
    : positive_(positive ? 1 : 0), use_linear_(use_linear) {
  // EnzoProlong.cpp:19-31
  if (type == "3A")
    method_ = 0;
  else if (type == "2A")
    method_ = 1;
  else if (type == "2B")
    method_ = 2;
  else
    ENZOE_ERROR("EnzoProlong::EnzoProlong", "Unrecognized interpolation method " + type);
}

void EnzoProlong::apply(precision_type precision, void *values_f, int m3_f[3],
                        int o3_f[3], int n3_f[3], const void *values_c, int m3_c[3],
                        int o3_c[3], int n3_c[3], bool accumulate) {
  ProlongLinear prolong_linear;
  if (!accumulate && !use_linear_) {
    // EnzoProlong.cpp:64-74 -> apply_ (:112-190), interpolate() on the device
    const int rank = (m3_f[1] == 1) ? 1 : ((m3_f[2] == 1) ? 2 : 3);
    ENZOE_ASSERT("EnzoProlong::apply()",
                 "Enzo's interpolate() is provided for rank 3 only", rank == 3);
    ENZOE_ASSERT("EnzoProlong::apply()",
                 "Unknown precision " + std::to_string(precision) +
                     " (double precision only)",
                 precision == precision_double || precision == precision_default);
    check_rc(enzoe_enzo_prolong(method_, positive_, static_cast<double *>(values_f), m3_f,
                                o3_f, n3_f, static_cast<const double *>(values_c), m3_c,
                                o3_c, n3_c, nullptr, g_stream),
             "EnzoProlong::apply_");
  } else if (!accumulate) {
    // EnzoProlong.cpp:62-78: the coarse region without its padding
    const int rank = (m3_f[1] == 1) ? 1 : ((m3_f[2] == 1) ? 2 : 3);
    int n3[3] = {n3_c[0], n3_c[1], n3_c[2]}, o3[3] = {o3_c[0], o3_c[1], o3_c[2]};
    for (int i = 0; i < rank; i++) {
      n3[i] -= 2;
      o3[i] += 1;
    }
    prolong_linear.apply(precision, values_f, m3_f, o3_f, n3_f, values_c, m3_c, o3, n3,
                         accumulate);
  } else {
    prolong_linear.apply(precision, values_f, m3_f, o3_f, n3_f, values_c, m3_c, o3_c, n3_c,
                         accumulate);
  }
}

int RestrictLinear::apply(precision_type precision, void *values_c, int nd3_c[3],
                          int im3_c[3], int n3_c[3], const void *values_f, int nd3_f[3],
                          int im3_f[3], int n3_f[3], bool accumulate) {
  ENZOE_ASSERT("RestrictLinear::apply()",
               "Unknown precision " + std::to_string(precision) +
                   " (double precision only)",
               precision == precision_double || precision == precision_default);
  check_rc(enzoe_restrict_linear(static_cast<double *>(values_c), nd3_c, im3_c, n3_c,
                                 static_cast<const double *>(values_f), nd3_f, im3_f, n3_f,
                                 accumulate ? 1 : 0, g_stream),
           "RestrictLinear::apply_");
  return (int)(sizeof(double) * n3_c[0] * n3_c[1] * n3_c[2]);
}

// ------------------------------------------------------- flux correction

namespace {

// parse_min_digit_map_ (problem_MethodFluxCorrect.cpp:15-52): a float is the
// density's; a list "[field, digits, ...]" pairs fields with digits
std::map<std::string, double> parse_min_digit_map(const ParameterGroup &p) {
  const std::string *v = p.param("min_digits");
  std::map<std::string, double> out;
  if (v == nullptr) return out;
  std::string t = *v;
  if (t.find('[') == std::string::npos) {
    out["density"] = p.value_float("min_digits", 0.0);
    return out;
  }
  for (char &c : t)
    if (c == '[' || c == ']' || c == ',' || c == '"') c = ' ';
  std::istringstream is(t);
  std::vector<std::string> items;
  std::string w;
  while (is >> w) items.push_back(w);
  ENZOE_ASSERT("parse_min_digit_map_",
               "The list assigned to " + p.full_name("min_digits") +
                   " must have a non-negative, even length",
               items.size() % 2 == 0);
  for (std::size_t i = 0; i < items.size(); i += 2) {
    ENZOE_ASSERT("parse_min_digit_map_",
                 "The \"" + items[i] + "\" field is listed more than once in " +
                     p.full_name("min_digits"),
                 out.find(items[i]) == out.end());
    out[items[i]] = std::atof(items[i + 1].c_str());
  }
  return out;
}

bool parse_logical(const ParameterGroup &p, const char *name, bool dflt) {
  const std::string v = lowered(p.value_string(name, dflt ? "true" : "false"));
  ENZOE_ASSERT("MethodFluxCorrect", p.full_name(name) + " must be true or false",
               v == "true" || v == "false");
  return v == "true";
}

}  // namespace

MethodFluxCorrect::MethodFluxCorrect(ParameterGroup p)
    : MethodFluxCorrect(p.value_string("group", "conserved"),
                        parse_logical(p, "enable", true), parse_min_digit_map(p)) {}

MethodFluxCorrect::MethodFluxCorrect(std::string group, bool enable,
                                     std::map<std::string, double> min_digits_map)
    : group_(std::move(group)), enable_(enable), min_digits_map_(std::move(min_digits_map)) {
  // problem_MethodFluxCorrect.cpp:67-112
  Grouping *groups = cello::field_groups();
  fields_ = groups->group_list(group_);
  ENZOE_ASSERT("MethodFluxCorrect::MethodFluxCorrect",
               "Must add field(s) to the field group \"" + group_ + "\"", !fields_.empty());
  for (const auto &kv : min_digits_map_) {
    ENZOE_ASSERT("MethodFluxCorrect::MethodFluxCorrect",
                 "Can't check conserved digits for the \"" + kv.first +
                     "\" field since it's not in the field group \"" + group_ + "\".",
                 groups->is_in(kv.first, group_));
  }
  field_sum_.assign(fields_.size(), 0.0L);
  field_sum_0_.assign(fields_.size(), 0.0L);
}

void MethodFluxCorrect::flux_correct_(Block *block) {
  FluxData *fd = block->flux_data();
  const int nf = fd->num_fields();
  if (!(enable_ && block->is_leaf()) || nf == 0) return;
  ENZOE_ASSERT("MethodFluxCorrect::flux_correct_", "the block holds no FluxData",
               fd->allocated());
  // faces whose neighbour is finer (:439-451)
  int correct[6];
  for (int axis = 0; axis < 3; axis++)
    for (int face = 0; face < 2; face++)
      correct[2 * axis + face] = block->face_level(axis, face) > block->level() ? 1 : 0;
  Grouping *groups = cello::field_groups();
  std::vector<double *> fields;
  std::vector<int> conservative;
  int i_density = -1;
  for (int i_f = 0; i_f < nf; i_f++) {
    const std::string &name = fd->field_name(i_f);
    if (name == "density") i_density = i_f;
    fields.push_back(block->field_view(name).data());
    conservative.push_back(groups->is_in(name, "make_field_conservative") ? 1 : 0);
  }
  for (int i_f = 0; i_f < nf; i_f++)
    ENZOE_ASSERT("MethodFluxCorrect::flux_correct_",
                 "The \"density\" field must exist to perform flux corrections on \"" +
                     fd->field_name(i_f) + "\".",
                 !(conservative[i_f] && i_f != i_density && i_density < 0));
  const std::array<int, 3> n = fd->size();
  check_rc(enzoe_flux_correct(fields.data(), nf, i_density, conservative.data(),
                              fd->flux_array(), n[0], n[1], n[2], block->ghost_depth(),
                              correct, g_stream),
           "MethodFluxCorrect::flux_correct_");
}

void MethodFluxCorrect::local_sums_(Block *block, std::vector<long double> &reduce) const {
  // compute_continue_refresh (:134-208): active-region sums, times density
  // for "make_field_conservative" fields, divided by 2^(level * rank)
  const EFlt3DArray &rho_v = block->field_view("density");
  const int mz = rho_v.shape(0), my = rho_v.shape(1), mx = rho_v.shape(2);
  const int g = block->ghost_depth();
  const std::size_t cells = (std::size_t)mx * my * mz;
  std::vector<double> rho(cells), values(cells);
  if (hipStreamSynchronize((hipStream_t)g_stream) != hipSuccess ||
      hipMemcpy(rho.data(), rho_v.data(), cells * sizeof(double), hipMemcpyDeviceToHost) !=
          hipSuccess)
    ENZOE_ERROR("MethodFluxCorrect::compute_continue_refresh", "hipMemcpy failed");
  Grouping *groups = cello::field_groups();
  const long double w = (long double)(1 << (block->level() * 3));
  for (std::size_t i_f = 0; i_f < fields_.size(); i_f++) {
    const EFlt3DArray a = block->field_view(fields_[i_f]);
    if (hipMemcpy(values.data(), a.data(), cells * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess)
      ENZOE_ERROR("MethodFluxCorrect::compute_continue_refresh", "hipMemcpy failed");
    const bool scale = groups->is_in(fields_[i_f], "make_field_conservative");
    long double sum = 0.0L;
    for (int iz = g; iz < mz - g; iz++)
      for (int iy = g; iy < my - g; iy++)
        for (int ix = g; ix < mx - g; ix++) {
          const std::size_t i = ix + (std::size_t)mx * (iy + (std::size_t)my * iz);
          if (scale)
            sum += values[i] * rho[i];
          else
            sum += values[i];
        }
    reduce[i_f] += sum / w;
  }
}

void MethodFluxCorrect::compute_blocks(const std::vector<Block *> &blocks) throw() {
  std::vector<long double> reduce(fields_.size(), 0.0L);
  for (Block *b : blocks) {
    flux_correct_(b);
    if (b->is_leaf()) local_sums_(b, reduce);
  }
  // compute_continue_sum_fields (:212-262)
  field_sum_ = reduce;
  if (cycle_ == 0) field_sum_0_ = field_sum_;
  min_digits_ok_ = true;
  for (std::size_t i_f = 0; i_f < fields_.size(); i_f++) {
    const long double a = field_sum_0_[i_f], b = field_sum_[i_f];
    const long double err = (a != 0.0L) ? std::fabs((a - b) / a) : std::fabs(a - b);
    const double digits = -std::log10((double)err);
    digits_[fields_[i_f]] = digits;
    auto it = min_digits_map_.find(fields_[i_f]);
    if (it != min_digits_map_.end() && !(digits >= it->second)) min_digits_ok_ = false;
  }
  cycle_++;
  for (Block *b : blocks) {
    b->flux_data()->deallocate();
    b->compute_done();
  }
}

void MethodFluxCorrect::compute(Block *block) throw() { compute_blocks({block}); }

double MethodFluxCorrect::conserved_digits(const std::string &field) const {
  auto it = digits_.find(field);
  return it == digits_.end() ? 0.0 : it->second;
}

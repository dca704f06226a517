// HDF5 block output and restart input in the layout of Cello's data files.
//
//   writer: OutputData::open / write_hierarchy / write_block /
//           write_field_data / close  (src/Cello/io_OutputData.cpp:68-250),
//           Output::expand_name_ (src/Cello/io_Output.cpp:125-207),
//           FileHdf5 attribute and dataset creation
//           (src/Cello/disk_FileHdf5.cpp:107-135, 236-290, 618-712, 807-880)
//   reader: IoEnzoReader::file_read_hierarchy_ / file_read_block_fields_
//           (src/Enzo/io/control_restart.cpp:590-720)
//   names:  Block::name (src/Cello/mesh_Block.cpp:798-821),
//           Index::bit_string (src/Cello/mesh_Index.cpp:544-575)
//
// Device-resident fields reach the file in slabs of whole planes (32 MiB at
// most) through two pinned host buffers: the copy of slab j+1 runs on the
// caller's stream while slab j is written, so the device-to-host transfer
// hides behind the file write.  The buffers are pinned once per process.
#include "../../include/enzoe_output.h"

#include <hdf5.h>
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <mutex>
#include <string>
#include <vector>

namespace {

// CELLO_VERSION of the files whose layout these are (CMakeLists.txt:20,78)
const char *const kLayoutVersion = "1.0.0";

constexpr int kMaxLevel = 20;    // INDEX_BITS_TREE (mesh_Index.hpp:9)
constexpr int kMaxArray = 1024;  // 1 << INDEX_BITS_ARRAY (mesh_Index.hpp:8)

bool index_ok(const enzoe_block_index *ix, const int *rb) {
  if (!ix || !rb) return false;
  if (ix->level < 0 || ix->level > kMaxLevel) return false;
  for (int a = 0; a < 3; ++a) {
    if (rb[a] < 1 || rb[a] > kMaxArray) return false;
    const long long n = (long long)rb[a] << ix->level;
    if (ix->global[a] < 0 || ix->global[a] >= n) return false;
  }
  return true;
}

// bits needed for root_blocks - 1 (mesh_Block.cpp:812-818)
int bit_length(int v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

std::string block_name(const enzoe_block_index &ix, const int rb[3], int rank) {
  const int level = ix.level;
  std::string s = "B";
  for (int a = 0; a < rank; ++a) {
    const int root = ix.global[a] >> level;
    for (int i = bit_length(rb[a] - 1) - 1; i >= 0; --i)
      s += ((root >> i) & 1) ? '1' : '0';
    if (level > 0) {
      s += ':';
      // child of level l: bit (level - l) of the position inside the root block
      for (int l = 1; l <= level; ++l) s += ((ix.global[a] >> (level - l)) & 1) ? '1' : '0';
    }
    if (a < rank - 1) s += '_';
  }
  return s;
}

// NodeBits {level:2, array:10, tree:20}, first member in the low bits
void index_words(const enzoe_block_index &ix, int w[3]) {
  const unsigned level = (unsigned)ix.level;
  // Index::set_level: the level's bits 0-1, 2-3 and 4 in the three level
  // fields (the last one's bit 0 is the sign, 0 here)
  const unsigned lev[3] = {level & 3u, (level >> 2) & 3u, (level >> 3) & 2u};
  for (int a = 0; a < 3; ++a) {
    const unsigned g = (unsigned)ix.global[a];
    const unsigned array = g >> level;
    // Index::set_child: the child of level l at tree bit (20 - l)
    const unsigned tree = level ? (g & ((1u << level) - 1u)) << (kMaxLevel - level) : 0u;
    const unsigned v = lev[a] | (array << 2) | (tree << 12);
    std::memcpy(&w[a], &v, sizeof v);
  }
}

// the printf conversion of one variable: flags, width, precision, no
// length modifier (an 'l' is accepted for a double), then the conversion
bool conversion_ok(const std::string &spec, bool is_double) {
  size_t i = 1;  // after '%'
  while (i < spec.size() && std::strchr("-+ #0", spec[i])) ++i;
  while (i < spec.size() && spec[i] >= '0' && spec[i] <= '9') ++i;
  if (i < spec.size() && spec[i] == '.') {
    ++i;
    while (i < spec.size() && spec[i] >= '0' && spec[i] <= '9') ++i;
  }
  if (is_double && i < spec.size() && spec[i] == 'l') ++i;
  if (i >= spec.size()) return false;
  const char c = spec[i];
  return is_double ? (std::strchr("fFeEgGaA", c) != nullptr)
                   : (std::strchr("diouxXc", c) != nullptr);
}

int expand_name(const char *format, const char *const *args, int nargs, int cycle,
                double time, int count, int proc, std::string *out) {
  if (!format || !*format) {
    out->clear();
    return ENZOE_OK;
  }
  const std::string name(format);
  if (nargs < 0 || (nargs > 0 && !args)) return ENZOE_ERR_ARG;
  if (name.find("\\%") != std::string::npos) return ENZOE_ERR_CONFIG;
  int npct = 0;
  for (char c : name) npct += (c == '%');
  if (npct != nargs) return ENZOE_ERR_CONFIG;
  std::string left = name, right;
  constexpr int kBuffer = 256;  // MAX_BUFFER + 1 (io_Output.cpp:135-138)
  char buffer[kBuffer];
  for (int i = 0; i < nargs; ++i) {
    const char *arg = args[nargs - 1 - i];  // right to left
    if (!arg) return ENZOE_ERR_ARG;
    const size_t pos = left.rfind('%');
    std::string middle = left.substr(pos);
    left = left.substr(0, pos);
    if (middle.size() >= (size_t)kBuffer) middle.resize(kBuffer - 1);
    const std::string a(arg);
    int n;
    if (a == "time") {
      if (!conversion_ok(middle, true)) return ENZOE_ERR_CONFIG;
      n = std::snprintf(buffer, sizeof buffer, middle.c_str(), time);
    } else {
      int value;
      if (a == "cycle") value = cycle;
      else if (a == "count") value = count;
      else if (a == "proc") value = proc;
      else if (a == "flipflop") value = count % 2;
      else return ENZOE_ERR_CONFIG;
      if (!conversion_ok(middle, false)) return ENZOE_ERR_CONFIG;
      n = std::snprintf(buffer, sizeof buffer, middle.c_str(), value);
    }
    if (n < 0) return ENZOE_ERR_CONFIG;
    right = std::string(buffer) + right;
  }
  *out = left + right;
  return ENZOE_OK;
}

int copy_out(const std::string &s, char *out, int out_len) {
  if (!out || out_len <= 0) return ENZOE_ERR_ARG;
  if ((int)s.size() + 1 > out_len) return ENZOE_ERR_ARG;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return ENZOE_OK;
}

// ------------------------------------------------------------------ HDF5

void quiet_hdf5() {
  static bool done = false;
  if (!done) {
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors are returned
    done = true;
  }
}

// FileHdf5::write_meta_: a one-dimensional attribute of n elements
bool write_attr(hid_t loc, const char *name, hid_t type, hsize_t n, const void *buf) {
  const hid_t space = H5Screate_simple(1, &n, nullptr);
  if (space < 0) return false;
  const hid_t attr = H5Acreate2(loc, name, type, space, H5P_DEFAULT, H5P_DEFAULT);
  bool ok = attr >= 0 && H5Awrite(attr, type, buf) >= 0;
  if (attr >= 0) ok = (H5Aclose(attr) >= 0) && ok;
  H5Sclose(space);
  return ok;
}

// reads an attribute of exactly n elements
bool read_attr(hid_t loc, const char *name, hid_t type, hsize_t n, void *buf) {
  const hid_t attr = H5Aopen(loc, name, H5P_DEFAULT);
  if (attr < 0) return false;
  const hid_t space = H5Aget_space(attr);
  bool ok = space >= 0 && H5Sget_simple_extent_npoints(space) == (hssize_t)n &&
            H5Aread(attr, type, buf) >= 0;
  if (space >= 0) H5Sclose(space);
  H5Aclose(attr);
  return ok;
}

// the dataset extents write_field_data gives a field (io_OutputData.cpp:
// 240-246): (mz, my, mx), or (my, mx) / (mx) when the higher extents are 1
int field_dims(const enzoe_output_field &f, hsize_t dims[3]) {
  if (f.mz > 1) {
    dims[0] = f.mz;
    dims[1] = f.my;
    dims[2] = f.mx;
    return 3;
  }
  if (f.my > 1) {
    dims[0] = f.my;
    dims[1] = f.mx;
    return 2;
  }
  dims[0] = f.mx;
  return 1;
}

bool field_ok(const enzoe_output_field &f) {
  return f.name && *f.name && f.data && f.mx >= 1 && f.my >= 1 && f.mz >= 1;
}

// one slab of one field: planes [z0, z0 + nz) of its slowest extent
struct Piece {
  int field;
  hsize_t z0, nz;
  size_t plane;  // bytes per plane
};

constexpr size_t kStageBytes = size_t(32) << 20;

// cuts every field into slabs of at most kStageBytes (one plane at least);
// *stage_bytes = the largest slab
std::vector<Piece> slabs(const enzoe_output_field *fields, int n, size_t *stage_bytes) {
  std::vector<Piece> out;
  size_t need = 0;
  for (int k = 0; k < n; ++k) {
    hsize_t dims[3];
    const int rank = field_dims(fields[k], dims);
    size_t plane = sizeof(double);
    for (int d = 1; d < rank; ++d) plane *= dims[d];
    const hsize_t per = std::max<hsize_t>(1, kStageBytes / plane);
    need = std::max(need, (size_t)std::min(per, dims[0]) * plane);
    for (hsize_t z0 = 0; z0 < dims[0]; z0 += per)
      out.push_back({k, z0, std::min(per, dims[0] - z0), plane});
  }
  *stage_bytes = need;
  return out;
}

// Two pinned host buffers shared by every call in the process (one call at
// a time, under the mutex), allocated on first use and kept until the
// process ends: pinning memory costs more than a slab's transfer.
struct PinnedPool {
  std::mutex mu;
  void *buf[2] = {nullptr, nullptr};
  size_t bytes = 0;

  int reserve(size_t need) {
    if (need <= bytes) return ENZOE_OK;
    for (auto &b : buf)
      if (b) {
        (void)hipHostFree(b);
        b = nullptr;
      }
    bytes = 0;
    for (auto &b : buf)
      if (hipHostMalloc(&b, need, hipHostMallocDefault) != hipSuccess) {
        b = nullptr;
        return ENZOE_ERR_ALLOC;
      }
    bytes = need;
    return ENZOE_OK;
  }
};

PinnedPool &pinned_pool() {
  static PinnedPool *pool = new PinnedPool;  // never destroyed (see above)
  return *pool;
}

struct Events {
  hipEvent_t e[2] = {nullptr, nullptr};
  int create() {
    for (auto &x : e)
      if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) return ENZOE_ERR_HIP;
    return ENZOE_OK;
  }
  ~Events() {
    for (auto &x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// selects the slab in the dataset's space and makes the matching memory space
hid_t slab_spaces(hid_t fspace, int rank, const hsize_t dims[3], const Piece &p) {
  hsize_t start[3] = {p.z0, 0, 0}, count[3] = {p.nz, 1, 1};
  for (int d = 1; d < rank; ++d) count[d] = dims[d];
  if (H5Sselect_hyperslab(fspace, H5S_SELECT_SET, start, nullptr, count, nullptr) < 0)
    return -1;
  const hsize_t n = p.nz * (p.plane / sizeof(double));
  return H5Screate_simple(1, &n, nullptr);
}

}  // namespace

struct enzoe_output_s {
  hid_t file = -1;
  std::string dir, file_name, path, list_base;
  FILE *block_list = nullptr;
  enzoe_output_hierarchy h{};
};

extern "C" {

int enzoe_block_name(const enzoe_block_index *index, const int root_blocks[3], int rank,
                     char *out, int out_len) {
  if (!index_ok(index, root_blocks) || rank < 1 || rank > 3) return ENZOE_ERR_ARG;
  return copy_out(block_name(*index, root_blocks, rank), out, out_len);
}

int enzoe_block_index_words(const enzoe_block_index *index, const int root_blocks[3],
                            int words[3]) {
  if (!index_ok(index, root_blocks) || !words) return ENZOE_ERR_ARG;
  index_words(*index, words);
  return ENZOE_OK;
}

int enzoe_output_expand_name(const char *format, const char *const *args, int nargs,
                             int cycle, double time, int count, int proc, char *out,
                             int out_len) {
  std::string s;
  const int rc = expand_name(format, args, nargs, cycle, time, count, proc, &s);
  return rc ? rc : copy_out(s, out, out_len);
}

int enzoe_output_open(const char *file_format, const char *const *file_args,
                      int n_file_args, const char *dir_format,
                      const char *const *dir_args, int n_dir_args, int count, int proc,
                      const enzoe_output_hierarchy *h, enzoe_output_t *out) {
  if (!out || !h || !file_format) return ENZOE_ERR_ARG;
  *out = nullptr;
  if (h->rank < 1 || h->rank > 3 || h->max_level < 0) return ENZOE_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (h->root_blocks[a] < 1 || h->root_blocks[a] > kMaxArray) return ENZOE_ERR_ARG;
  std::string file_name, name_dir;
  int rc = expand_name(file_format, file_args, n_file_args, h->cycle, h->time, count, proc,
                       &file_name);
  if (rc) return rc;
  rc = expand_name(dir_format, dir_args, n_dir_args, h->cycle, h->time, count, proc,
                   &name_dir);
  if (rc) return rc;
  if (file_name.empty()) return ENZOE_ERR_CONFIG;
  quiet_hdf5();
  auto *o = new enzoe_output_s;
  o->h = *h;
  o->file_name = file_name;
  // Output::directory (io_Output.hpp:233-243) and OutputData::write_block's
  // list names (io_OutputData.cpp:138-152)
  if (name_dir.empty()) {
    o->dir = ".";
    o->list_base = file_name.substr(0, file_name.rfind('.'));
  } else {
    o->dir = name_dir;
    o->list_base = name_dir;
    std::error_code ec;
    std::filesystem::create_directories(name_dir, ec);
    if (ec && !std::filesystem::is_directory(name_dir)) {
      delete o;
      return ENZOE_ERR_IO;
    }
  }
  o->path = o->dir + "/" + file_name;
  o->file = H5Fcreate(o->path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (o->file < 0) {
    delete o;
    return ENZOE_ERR_IO;
  }
  // OutputData::write_hierarchy (io_OutputData.cpp:106-122)
  bool ok = write_attr(o->file, "rank", H5T_NATIVE_INT, 1, &h->rank) &&
            write_attr(o->file, "cycle", H5T_NATIVE_INT, 1, &h->cycle) &&
            write_attr(o->file, "time", H5T_NATIVE_DOUBLE, 1, &h->time) &&
            write_attr(o->file, "dt", H5T_NATIVE_DOUBLE, 1, &h->dt) &&
            write_attr(o->file, "lower", H5T_NATIVE_DOUBLE, 3, h->lower) &&
            write_attr(o->file, "upper", H5T_NATIVE_DOUBLE, 3, h->upper) &&
            write_attr(o->file, "max_level", H5T_NATIVE_INT, 1, &h->max_level) &&
            write_attr(o->file, "version", H5T_NATIVE_CHAR, std::strlen(kLayoutVersion),
                       kLayoutVersion);
  const std::string list = o->dir + "/" + o->list_base;
  FILE *fl = ok ? std::fopen((list + ".file_list").c_str(), "w") : nullptr;
  if (fl) {
    ok = std::fprintf(fl, "%s\n", file_name.c_str()) > 0;
    ok = (std::fclose(fl) == 0) && ok;
  } else {
    ok = false;
  }
  if (ok) o->block_list = std::fopen((list + ".block_list").c_str(), "w");
  if (!ok || !o->block_list) {
    enzoe_output_close(o);
    return ENZOE_ERR_IO;
  }
  *out = o;
  return ENZOE_OK;
}

const char *enzoe_output_file_path(enzoe_output_t o) { return o ? o->path.c_str() : nullptr; }

int enzoe_output_write_block(enzoe_output_t o, const enzoe_output_block *b,
                             enzoe_stream_t stream) {
  if (!o || !b || o->file < 0) return ENZOE_ERR_ARG;
  const enzoe_output_hierarchy &h = o->h;
  if (!index_ok(&b->index, h.root_blocks) || b->nfields < 0 || (b->nfields && !b->fields))
    return ENZOE_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (b->n[a] < 1 || b->ghost[a] < 0) return ENZOE_ERR_ARG;
  for (int k = 0; k < b->nfields; ++k)
    if (!field_ok(b->fields[k])) return ENZOE_ERR_ARG;
  const std::string name = block_name(b->index, h.root_blocks, h.rank);
  const hid_t group =
      H5Gcreate2(o->file, ("/" + name).c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (group < 0) return ENZOE_ERR_IO;

  // IoBlock::set_block / IoEnzoBlock::set_block: Block::lower/upper
  // (mesh_Block.cpp:832-886), EnzoBlock::initialize (EnzoBlock.cpp:170-245)
  const int level = b->index.level;
  int words[3], array[3], dim[3], start[3], end[3];
  double lower[3], upper[3], width[3];
  index_words(b->index, words);
  for (int a = 0; a < 3; ++a) {
    const int nb = h.root_blocks[a] << level;
    const double am = 1.0 * b->index.global[a] / nb;
    const double ap = 1.0 * (b->index.global[a] + 1) / nb;
    lower[a] = (1.0 - am) * h.lower[a] + am * h.upper[a];
    upper[a] = (1.0 - ap) * h.lower[a] + ap * h.upper[a];
    width[a] = (upper[a] - lower[a]) / b->n[a];  // FieldData::cell_width
    array[a] = b->index.global[a] >> level;
    dim[a] = b->n[a] + 2 * b->ghost[a];
    start[a] = b->ghost[a];
    end[a] = b->ghost[a] + b->n[a] - 1;
  }
  const int num_field_data = 1;
  const long long order = 0;  // no MethodOrder* has run
  const double redshift = 0.0;
  bool ok = write_attr(group, "num_field_data", H5T_NATIVE_INT, 1, &num_field_data) &&
            write_attr(group, "index", H5T_NATIVE_INT, 3, words) &&
            write_attr(group, "lower", H5T_NATIVE_DOUBLE, 3, lower) &&
            write_attr(group, "upper", H5T_NATIVE_DOUBLE, 3, upper) &&
            write_attr(group, "cycle", H5T_NATIVE_INT, 1, &b->cycle) &&
            write_attr(group, "time", H5T_NATIVE_DOUBLE, 1, &b->time) &&
            write_attr(group, "dt", H5T_NATIVE_DOUBLE, 1, &b->dt) &&
            write_attr(group, "array", H5T_NATIVE_INT, 3, array) &&
            write_attr(group, "index_order", H5T_NATIVE_LLONG, 1, &order) &&
            write_attr(group, "count_order", H5T_NATIVE_LLONG, 1, &order) &&
            write_attr(group, "enzo_dt", H5T_NATIVE_DOUBLE, 1, &b->dt) &&
            write_attr(group, "enzo_GridLeftEdge", H5T_NATIVE_DOUBLE, 3, lower) &&
            write_attr(group, "enzo_GridDimension", H5T_NATIVE_INT, 3, dim) &&
            write_attr(group, "enzo_GridStartIndex", H5T_NATIVE_INT, 3, start) &&
            write_attr(group, "enzo_GridEndIndex", H5T_NATIVE_INT, 3, end) &&
            write_attr(group, "enzo_CellWidth", H5T_NATIVE_DOUBLE, 3, width) &&
            write_attr(group, "enzo_redshift", H5T_NATIVE_DOUBLE, 1, &redshift);

  // write_field_data, one dataset per field, written slab by slab; from
  // device memory the copy of slab j+1 runs while slab j is written
  int rc = ok ? ENZOE_OK : ENZOE_ERR_IO;
  size_t stage = 0;
  const std::vector<Piece> pieces = slabs(b->fields, b->nfields, &stage);
  PinnedPool &pool = pinned_pool();
  std::unique_lock<std::mutex> lock(pool.mu, std::defer_lock);
  Events ev;
  hipStream_t s = static_cast<hipStream_t>(stream);
  bool queued = false;
  auto issue = [&](size_t j) -> int {
    const Piece &p = pieces[j];
    const char *src = static_cast<const char *>(static_cast<void *>(b->fields[p.field].data));
    if (hipMemcpyAsync(pool.buf[j & 1], src + p.z0 * p.plane, p.nz * p.plane,
                       hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(ev.e[j & 1], s) != hipSuccess)
      return ENZOE_ERR_HIP;
    queued = true;
    return ENZOE_OK;
  };
  if (!rc && b->on_device && !pieces.empty()) {
    lock.lock();
    rc = pool.reserve(stage);
    if (!rc) rc = ev.create();
    if (!rc) rc = issue(0);
  }
  hid_t ds = -1, fspace = -1;
  hsize_t dims[3];
  int rank = 0, cur = -1;
  auto close_dataset = [&]() {
    if (ds >= 0 && H5Dclose(ds) < 0 && !rc) rc = ENZOE_ERR_IO;
    if (fspace >= 0) H5Sclose(fspace);
    ds = fspace = -1;
  };
  for (size_t j = 0; !rc && j < pieces.size(); ++j) {
    const Piece &p = pieces[j];
    const enzoe_output_field &f = b->fields[p.field];
    if (p.field != cur) {
      close_dataset();
      rank = field_dims(f, dims);
      fspace = H5Screate_simple(rank, dims, nullptr);
      ds = fspace < 0 ? -1
                      : H5Dcreate2(group, (std::string("field_") + f.name).c_str(),
                                   H5T_NATIVE_DOUBLE, fspace, H5P_DEFAULT, H5P_DEFAULT,
                                   H5P_DEFAULT);
      if (ds < 0) {
        rc = ENZOE_ERR_IO;
        break;
      }
      cur = p.field;
    }
    const void *src = static_cast<const char *>(static_cast<void *>(f.data)) + p.z0 * p.plane;
    if (b->on_device) {
      if (j + 1 < pieces.size()) rc = issue(j + 1);
      if (!rc && hipEventSynchronize(ev.e[j & 1]) != hipSuccess) rc = ENZOE_ERR_HIP;
      if (rc) break;
      src = pool.buf[j & 1];
    }
    const hid_t mspace = slab_spaces(fspace, rank, dims, p);
    if (mspace < 0 || H5Dwrite(ds, H5T_NATIVE_DOUBLE, mspace, fspace, H5P_DEFAULT, src) < 0)
      rc = ENZOE_ERR_IO;
    if (mspace >= 0) H5Sclose(mspace);
  }
  close_dataset();
  // no copy may still target the staging buffers once this returns
  if (queued && hipStreamSynchronize(s) != hipSuccess && !rc) rc = ENZOE_ERR_HIP;
  if (H5Gclose(group) < 0 && !rc) rc = ENZOE_ERR_IO;
  if (!rc && std::fprintf(o->block_list, "%s %s\n", name.c_str(), o->file_name.c_str()) < 0)
    rc = ENZOE_ERR_IO;
  return rc;
}

int enzoe_output_close(enzoe_output_t o) {
  if (!o) return ENZOE_ERR_ARG;
  int rc = ENZOE_OK;
  if (o->block_list && std::fclose(o->block_list) != 0) rc = ENZOE_ERR_IO;
  if (o->file >= 0 && H5Fclose(o->file) < 0) rc = ENZOE_ERR_IO;
  delete o;
  return rc;
}

int enzoe_input_read_state(const char *path, int *cycle, double *time, double *dt) {
  if (!path || !cycle || !time || !dt) return ENZOE_ERR_ARG;
  quiet_hdf5();
  const hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return ENZOE_ERR_IO;
  const bool ok = read_attr(file, "cycle", H5T_NATIVE_INT, 1, cycle) &&
                  read_attr(file, "time", H5T_NATIVE_DOUBLE, 1, time) &&
                  read_attr(file, "dt", H5T_NATIVE_DOUBLE, 1, dt);
  H5Fclose(file);
  return ok ? ENZOE_OK : ENZOE_ERR_IO;
}

int enzoe_input_read_block(const char *path, const char *block_name,
                           const enzoe_output_field *fields, int nfields, int on_device,
                           enzoe_stream_t stream, int *cycle, double *time, double *dt) {
  if (!path || !block_name || nfields < 0 || (nfields && !fields)) return ENZOE_ERR_ARG;
  for (int k = 0; k < nfields; ++k)
    if (!field_ok(fields[k])) return ENZOE_ERR_ARG;
  quiet_hdf5();
  const hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return ENZOE_ERR_IO;
  const hid_t group = H5Gopen2(file, (std::string("/") + block_name).c_str(), H5P_DEFAULT);
  int rc = group < 0 ? ENZOE_ERR_IO : ENZOE_OK;
  if (!rc && ((cycle && !read_attr(group, "cycle", H5T_NATIVE_INT, 1, cycle)) ||
              (time && !read_attr(group, "time", H5T_NATIVE_DOUBLE, 1, time)) ||
              (dt && !read_attr(group, "dt", H5T_NATIVE_DOUBLE, 1, dt))))
    rc = ENZOE_ERR_IO;
  size_t stage = 0;
  const std::vector<Piece> pieces = slabs(fields, nfields, &stage);
  PinnedPool &pool = pinned_pool();
  std::unique_lock<std::mutex> lock(pool.mu, std::defer_lock);
  Events ev;
  hipStream_t s = static_cast<hipStream_t>(stream);
  bool queued = false;
  if (!rc && on_device && !pieces.empty()) {
    lock.lock();
    rc = pool.reserve(stage);
    if (!rc) rc = ev.create();
  }
  hid_t ds = -1, fspace = -1;
  hsize_t dims[3];
  int rank = 0, cur = -1;
  auto close_dataset = [&]() {
    if (fspace >= 0) H5Sclose(fspace);
    if (ds >= 0) H5Dclose(ds);
    ds = fspace = -1;
  };
  for (size_t j = 0; !rc && j < pieces.size(); ++j) {
    const Piece &p = pieces[j];
    const enzoe_output_field &f = fields[p.field];
    if (p.field != cur) {
      close_dataset();
      cur = p.field;
      ds = H5Dopen2(group, (std::string("field_") + f.name).c_str(), H5P_DEFAULT);
      rank = field_dims(f, dims);
      hsize_t got[3];
      fspace = ds < 0 ? -1 : H5Dget_space(ds);
      const hid_t type = ds < 0 ? -1 : H5Dget_type(ds);
      bool ok = fspace >= 0 && type >= 0 && H5Tget_class(type) == H5T_FLOAT &&
                H5Sget_simple_extent_ndims(fspace) == rank &&
                H5Sget_simple_extent_dims(fspace, got, nullptr) == rank;
      for (int d = 0; ok && d < rank; ++d) ok = got[d] == dims[d];
      if (type >= 0) H5Tclose(type);
      if (!ok) {
        rc = ENZOE_ERR_IO;
        break;
      }
    }
    char *dst = static_cast<char *>(static_cast<void *>(f.data)) + p.z0 * p.plane;
    void *into = dst;
    if (on_device) {
      // the copy that last read this buffer (slab j-2) must be done
      if (j >= 2 && hipEventSynchronize(ev.e[j & 1]) != hipSuccess) {
        rc = ENZOE_ERR_HIP;
        break;
      }
      into = pool.buf[j & 1];
    }
    const hid_t mspace = slab_spaces(fspace, rank, dims, p);
    if (mspace < 0 || H5Dread(ds, H5T_NATIVE_DOUBLE, mspace, fspace, H5P_DEFAULT, into) < 0)
      rc = ENZOE_ERR_IO;
    if (mspace >= 0) H5Sclose(mspace);
    if (!rc && on_device) {
      queued = true;
      if (hipMemcpyAsync(dst, into, p.nz * p.plane, hipMemcpyHostToDevice, s) != hipSuccess ||
          hipEventRecord(ev.e[j & 1], s) != hipSuccess)
        rc = ENZOE_ERR_HIP;
    }
  }
  close_dataset();
  if (queued && hipStreamSynchronize(s) != hipSuccess && !rc) rc = ENZOE_ERR_HIP;
  if (group >= 0) H5Gclose(group);
  H5Fclose(file);
  return rc;
}

}  // extern "C"

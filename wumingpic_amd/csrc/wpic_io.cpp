// wpic_io.cpp -- the reference's raw + JSON output format at one rank, on the
// host: paraio (2d/common/paraio.f90) over jsonio (utils/iocore/jsonio.f90)
// and mpiio (utils/iocore/mpiio.f90).
//
// A file pair <name>.raw / <name>.json: the raw file holds little-endian
// values back to back (attributes written by rank 0 at increasing offsets,
// mpiio.f90:272-294, then whole arrays in Fortran order, which is what the
// collective writes of mpiio.f90:440-663 leave at one rank); the JSON file
// describes each entry (datatype, offset, size, ndim, shape, description,
// and the value itself for attributes, jsonio.f90:128-170) and is printed the
// way JSON-Fortran's json%print prints it (two-space indentation, one array
// element per line, reals written with (ss,E27.17E4) and then compacted).
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wpic_io.h"
#include "wpic_json.h"

namespace wpic {
int set_error(int code, const std::string& msg);
}

namespace {

using wpic::set_error;
using namespace wpic::json;

// mpiio_get_endian_flag (mpiio.f90:101-108): the bytes 1,0,0,0 read as int32
int32_t endian_flag() {
  const unsigned char b[4] = {1, 0, 0, 0};
  int32_t f;
  std::memcpy(&f, b, 4);
  return f;
}

// ---------------------------------------------------------------------------
// One output file pair being written: the JSON tree and the raw stream at
// offset disp, as a paraio routine builds them.
// ---------------------------------------------------------------------------
struct Writer {
  JVal root = JVal::object();
  JVal* attr = nullptr;
  JVal* dset = nullptr;
  FILE* raw = nullptr;
  long long disp = 0;
  bool failed = false;

  bool open(const std::string& rawpath, const std::string& rawname) {
    raw = std::fopen(rawpath.c_str(), "wb");
    if (!raw) return false;
    JVal& meta = root.add("meta", JVal::object());
    meta.add("endian", JVal::integer(endian_flag()));
    meta.add("rawfile", JVal::string(rawname));
    return true;
  }
  void begin_attributes() { attr = &root.add("attribute", JVal::object()); }
  void begin_datasets() { dset = &root.add("dataset", JVal::object()); }

  void put_bytes(const void* data, size_t n) {
    if (n && std::fwrite(data, 1, n, raw) != n) failed = true;
    disp += (long long)n;
  }
  // jsonio.f90 put_metadata_ik: datatype, offset, size, ndim, shape, description
  static JVal metadata(const char* dtype, long long off, long long size, const std::vector<long long>& shape,
                       const std::string& desc) {
    JVal o = JVal::object();
    o.add("datatype", JVal::string(dtype));
    o.add("offset", JVal::integer(off));
    o.add("size", JVal::integer(size));
    o.add("ndim", JVal::integer((long long)shape.size()));
    JVal sh;
    sh.k = JVal::ARR;
    for (long long n : shape) sh.arr.push_back(JVal::integer(n));
    o.add("shape", sh);
    o.add("description", JVal::string(desc));
    return o;
  }
  // jsonio_put_attribute + mpiio_write_atomic of a scalar i4 / r8 / r8 array
  void attr_i4(const char* name, int32_t v) {
    JVal& o = attr->add(name, metadata("i4", disp, 4, {1}, ""));
    o.add("data", JVal::integer(v));
    put_bytes(&v, 4);
  }
  void attr_r8(const char* name, double v) {
    JVal& o = attr->add(name, metadata("f8", disp, 8, {1}, ""));
    o.add("data", JVal::real(v));
    put_bytes(&v, 8);
  }
  void attr_r8v(const char* name, const double* v, int n) {
    JVal& o = attr->add(name, metadata("f8", disp, 8LL * n, {n}, ""));
    JVal a;
    a.k = JVal::ARR;
    for (int k = 0; k < n; k++) a.arr.push_back(JVal::real(v[k]));
    o.add("data", a);
    put_bytes(v, 8 * (size_t)n);
  }
  // jsonio_put_metadata + mpiio_write_collective of a whole array
  void dataset(const char* name, const char* dtype, const std::vector<long long>& shape, const void* data, size_t elem,
               const std::string& desc) {
    long long n = 1;
    for (long long s : shape) n *= s;
    dset->add(name, metadata(dtype, disp, n * (long long)elem, shape, desc));
    put_bytes(data, (size_t)n * elem);
  }
  bool finish(const std::string& jsonpath) {
    if (raw && std::fclose(raw) != 0) failed = true;
    raw = nullptr;
    std::string txt;
    print(root, 0, txt);
    txt += "\n";
    FILE* f = std::fopen(jsonpath.c_str(), "wb");
    if (!f) return false;
    const bool w = std::fwrite(txt.data(), 1, txt.size(), f) == txt.size();
    return std::fclose(f) == 0 && w && !failed;
  }
  ~Writer() {
    if (raw) std::fclose(raw);
  }
};

// '(i7.7, a)' of an iteration count: seven digits, asterisks when it does
// not fit
std::string it7(int it) {
  if (it < 0 || it > 9999999) return "*******";
  char b[16];
  std::snprintf(b, sizeof b, "%07d", it);
  return b;
}

std::string two(int n) {
  char b[8];
  if (n < 0 || n > 99) return "**";
  std::snprintf(b, sizeof b, "%02d", n);
  return b;
}

}  // namespace

// ---------------------------------------------------------------------------
// the module state of paraio__init (2d/common/paraio.f90:37-78,
// 3d/common/paraio.f90:37-85)
// ---------------------------------------------------------------------------
struct wpic_paraio {
  int dims;  // 2 or 3
  int ndim, np, nsp, nxgs, nxge, nygs, nyge, nzgs, nzge, nys, nye, nzs, nze;
  int nproc;
  double delx, delt, c;
  std::vector<double> q, r;
  std::string dir;
};

namespace {

int nx_of(const wpic_paraio* io) { return io->nxge - io->nxgs + 1; }
int ny_of(const wpic_paraio* io) { return io->nye - io->nys + 1; }
int nz_of(const wpic_paraio* io) { return io->nze - io->nzs + 1; }

// put_metadata (2d/common/paraio.f90:956-1002; 3d/common/paraio.f90:1087-1139)
void put_common(Writer& w, const wpic_paraio* io) {
  w.attr_i4("ndim", io->ndim);
  w.attr_i4("np", io->np);
  w.attr_i4("nxgs", io->nxgs);
  w.attr_i4("nxge", io->nxge);
  w.attr_i4("nygs", io->nygs);
  w.attr_i4("nyge", io->nyge);
  if (io->dims == 3) {
    w.attr_i4("nzgs", io->nzgs);
    w.attr_i4("nzge", io->nzge);
  }
  w.attr_i4("nsp", io->nsp);
  w.attr_i4("nproc", io->nproc);
  w.attr_r8("delx", io->delx);
  w.attr_r8("delt", io->delt);
  w.attr_r8("c", io->c);
  w.attr_r8v("r", io->r.data(), io->nsp);
  w.attr_r8v("q", io->q.data(), io->nsp);
}

// the dummy attribute every file starts with: MPI_OFFSET_KIND as a default
// integer (paraio.f90:18, 140)
constexpr int32_t MOK = 8;

int open_writer(Writer& w, const wpic_paraio* io, const std::string& name) {
  if (!w.open(io->dir + name + ".raw", name + ".raw"))
    return set_error(WPIC_EINVAL, "cannot open " + io->dir + name + ".raw: " + std::strerror(errno));
  return 0;
}

int close_writer(Writer& w, const wpic_paraio* io, const std::string& name) {
  if (!w.finish(io->dir + name + ".json"))
    return set_error(WPIC_EINVAL, "cannot write " + io->dir + name + ".json or its raw file");
  return 0;
}

// column-major index helpers of the reference arrays
inline size_t i_uf2(const wpic_paraio* io, int c, int i, int j) {
  return (size_t)(c - 1) + 6 * ((size_t)(i - (io->nxgs - 2)) + (size_t)(nx_of(io) + 4) * (size_t)(j - (io->nys - 2)));
}
inline size_t i_mom2(const wpic_paraio* io, int k, int i, int j, int s) {
  return (size_t)(k - 1) +
         7 * ((size_t)(i - (io->nxgs - 1)) +
              (size_t)(nx_of(io) + 2) * ((size_t)(j - (io->nys - 1)) + (size_t)(ny_of(io) + 2) * (size_t)(s - 1)));
}
inline size_t i_uf3(const wpic_paraio* io, int c, int i, int j, int k) {
  return (size_t)(c - 1) +
         6 * ((size_t)(i - (io->nxgs - 2)) +
              (size_t)(nx_of(io) + 4) * ((size_t)(j - (io->nys - 2)) + (size_t)(ny_of(io) + 4) * (size_t)(k - (io->nzs - 2))));
}
inline size_t i_mom3(const wpic_paraio* io, int l, int i, int j, int k, int s) {
  return (size_t)(l - 1) +
         7 * ((size_t)(i - (io->nxgs - 1)) +
              (size_t)(nx_of(io) + 2) *
                  ((size_t)(j - (io->nys - 1)) +
                   (size_t)(ny_of(io) + 2) * ((size_t)(k - (io->nzs - 1)) + (size_t)(nz_of(io) + 2) * (size_t)(s - 1))));
}

// the active particles, species by species, in storage order: all of them
// (mode 0) or those whose id, read as a 64-bit integer, is positive (mode 1)
// (get_particle_count, 2d/common/paraio.f90:897-938)
void pack_particles(const wpic_paraio* io, const double* up, const int* np2, int mode, std::vector<double>& buf,
                    std::vector<long long>& count) {
  const int nrow = io->dims == 3 ? ny_of(io) * nz_of(io) : ny_of(io);
  count.assign(io->nsp, 0);
  buf.clear();
  for (int s = 0; s < io->nsp; s++)
    for (int rw = 0; rw < nrow; rw++) {
      const int n = np2[(size_t)rw + (size_t)nrow * s];
      for (int ii = 0; ii < n; ii++) {
        const double* u = up + (size_t)io->ndim * ((size_t)ii + (size_t)io->np * ((size_t)rw + (size_t)nrow * s));
        if (mode == 1) {
          int64_t pid;
          std::memcpy(&pid, &u[io->ndim - 1], 8);
          if (pid <= 0) continue;
        }
        count[s]++;
        buf.insert(buf.end(), u, u + io->ndim);
      }
    }
}

int write_particles(wpic_paraio* io, const double* up, const double* uf, const int* np2, int it, int mode,
                    const char* suffix) {
  const std::string name = it7(it) + suffix;
  Writer w;
  int rc = open_writer(w, io, name);
  if (rc) return rc;
  w.begin_attributes();
  w.attr_i4("dummy_attribute", MOK);
  w.attr_i4("it", it);
  put_common(w, io);
  w.begin_datasets();
  const int nx = nx_of(io), ny = ny_of(io);
  if (io->dims == 2) {
    // uf(1:6, nxgs:nxge, nys:nye) (paraio.f90:805-818)
    std::vector<double> f((size_t)6 * nx * ny);
    for (int j = io->nys; j <= io->nye; j++)
      for (int i = io->nxgs; i <= io->nxge; i++)
        for (int c = 1; c <= 6; c++)
          f[(size_t)(c - 1) + 6 * ((size_t)(i - io->nxgs) + (size_t)nx * (j - io->nys))] = uf[i_uf2(io, c, i, j)];
    w.dataset("uf", "f8", {6, nx, ny}, f.data(), 8, "electromagnetic field");
  } else {
    const int nz = nz_of(io);
    std::vector<double> f((size_t)6 * nx * ny * nz);
    for (int k = io->nzs; k <= io->nze; k++)
      for (int j = io->nys; j <= io->nye; j++)
        for (int i = io->nxgs; i <= io->nxge; i++)
          for (int c = 1; c <= 6; c++)
            f[(size_t)(c - 1) + 6 * ((size_t)(i - io->nxgs) + (size_t)nx * ((size_t)(j - io->nys) + (size_t)ny * (k - io->nzs)))] =
                uf[i_uf3(io, c, i, j, k)];
    w.dataset("uf", "f8", {6, nx, ny, nz}, f.data(), 8, "electromagnetic field");
  }
  std::vector<double> buf;
  std::vector<long long> count;
  pack_particles(io, up, np2, mode, buf, count);
  size_t off = 0;
  for (int s = 0; s < io->nsp; s++) {
    const std::string nm = "up" + two(s + 1), desc = "particle species #" + two(s + 1);
    w.dataset(nm.c_str(), "f8", {io->ndim, count[s]}, buf.data() + off, 8, desc);
    off += (size_t)io->ndim * count[s];
  }
  return close_writer(w, io, name);
}

const JVal* meta_of(const JVal* grp, const char* name, long long& off, std::vector<long long>& shape) {
  const JVal* o = grp ? grp->get(name) : nullptr;
  if (!o) return nullptr;
  const JVal* po = o->get("offset");
  const JVal* sh = o->get("shape");
  if (!po || po->k != JVal::INT || !sh || sh->k != JVal::ARR) return nullptr;
  off = po->i;
  shape.clear();
  for (auto& v : sh->arr) shape.push_back(v.k == JVal::INT ? v.i : -1);
  return o;
}

bool read_at(FILE* f, long long off, void* dst, size_t n) {
  if (std::fseek(f, (long)off, SEEK_SET) != 0) return false;
  return std::fread(dst, 1, n, f) == n;
}

}  // namespace

extern "C" {

int wpic_paraio_init2d(wpic_paraio** out, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge, int nys,
                       int nye, int nproc, int nrank, double delx, double delt, double c, const double* q,
                       const double* r, const char* dir) {
  if (!out || !q || !r || !dir) return set_error(WPIC_EINVAL, "null argument");
  *out = nullptr;
  if (nproc != 1 || nrank != 0 || nys != nygs || nye != nyge)
    return set_error(WPIC_EINVAL, "paraio: single rank only (nproc = 1, nys = nygs, nye = nyge)");
  if (nsp < 1 || ndim < 1 || np < 0 || nxge < nxgs || nyge < nygs) return set_error(WPIC_EINVAL, "paraio: bad grid");
  wpic_paraio* io = new wpic_paraio();
  io->dims = 2;
  io->ndim = ndim;
  io->np = np;
  io->nsp = nsp;
  io->nxgs = nxgs;
  io->nxge = nxge;
  io->nygs = nygs;
  io->nyge = nyge;
  io->nys = nys;
  io->nye = nye;
  io->nzgs = io->nzge = io->nzs = io->nze = 0;
  io->nproc = nproc;
  io->delx = delx;
  io->delt = delt;
  io->c = c;
  io->q.assign(q, q + nsp);
  io->r.assign(r, r + nsp);
  io->dir = dir;
  *out = io;
  return 0;
}

int wpic_paraio_init3d(wpic_paraio** out, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge, int nzgs,
                       int nzge, int nys, int nye, int nzs, int nze, int nproc, int nrank, double delx, double delt,
                       double c, const double* q, const double* r, const char* dir) {
  int rc = wpic_paraio_init2d(out, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nproc, nrank, delx, delt, c, q, r,
                              dir);
  if (rc) return rc;
  if (nzs != nzgs || nze != nzge || nzge < nzgs) {
    delete *out;
    *out = nullptr;
    return set_error(WPIC_EINVAL, "paraio: single rank only (nzs = nzgs, nze = nzge)");
  }
  (*out)->dims = 3;
  (*out)->nzgs = nzgs;
  (*out)->nzge = nzge;
  (*out)->nzs = nzs;
  (*out)->nze = nze;
  return 0;
}

int wpic_paraio_finalize(wpic_paraio* io) {
  delete io;
  return 0;
}

// paraio__output: 2d/common/paraio.f90:93-228 (the padded up, np2 and uf with
// ghosts) or 3d/common/paraio.f90:102-283 (compact particles per species and
// their offsets, np2, uf with ghosts)
int wpic_paraio_output(wpic_paraio* io, const double* up, const double* uf, const int* np2, int nxs, int nxe, int it,
                       const char* filename) {
  if (!io) return set_error(WPIC_EINVAL, "Initialize first by calling paraio__init()");
  if (!up || !uf || !np2 || !filename) return set_error(WPIC_EINVAL, "null argument");
  const std::string name = filename;
  Writer w;
  int rc = open_writer(w, io, name);
  if (rc) return rc;
  w.begin_attributes();
  w.attr_i4("dummy_attribute", MOK);
  w.attr_i4("it", it);
  w.attr_i4("nxs", nxs);
  w.attr_i4("nxe", nxe);
  put_common(w, io);
  w.begin_datasets();
  const int nx = nx_of(io), ny = ny_of(io);
  if (io->dims == 2) {
    w.dataset("up", "f8", {io->ndim, io->np, ny, io->nsp, io->nproc}, up, 8, "particles");
    w.dataset("np2", "i4", {ny, io->nsp, io->nproc}, np2, 4, "number of active particles");
    w.dataset("uf", "f8", {6, nx + 4, ny + 4, io->nproc}, uf, 8, "electromagnetic fields including ghost cells");
  } else {
    const int nz = nz_of(io);
    w.dataset("np2", "i4", {ny, nz, io->nsp, io->nproc}, np2, 4, "number of active particles");
    std::vector<double> buf;
    std::vector<long long> count;
    pack_particles(io, up, np2, 0, buf, count);
    std::vector<long long> poffset(io->nsp, 0);  // the rank's offset in each species: 0 at one rank
    size_t off = 0;
    for (int s = 0; s < io->nsp; s++) {
      const std::string nm = "up" + two(s + 1), desc = "particle species #" + two(s + 1);
      w.dataset(nm.c_str(), "f8", {io->ndim, count[s]}, buf.data() + off, 8, desc);
      off += (size_t)io->ndim * count[s];
    }
    w.dataset("poffset", "i8", {io->nsp, io->nproc}, poffset.data(), 8, "particle offset");
    w.dataset("uf", "f8", {6, nx + 4, ny + 4, nz + 4, io->nproc}, uf, 8,
              "electromagnetic fields including ghost cells");
  }
  return close_writer(w, io, name);
}

// paraio__input: 2d/common/paraio.f90:233-426, 3d/common/paraio.f90:288-527
int wpic_paraio_input(wpic_paraio* io, double* up, double* uf, int* np2, int* nxs, int* nxe, int* it,
                      const char* filename) {
  if (!io) return set_error(WPIC_EINVAL, "Initialize first by calling paraio__init()");
  if (!up || !uf || !np2 || !nxs || !nxe || !it || !filename) return set_error(WPIC_EINVAL, "null argument");
  const std::string name = filename;
  std::string txt;
  if (!read_file(io->dir + name + ".json", txt)) return set_error(WPIC_EINVAL, "cannot read " + io->dir + name + ".json");
  Parser ps{txt.data(), txt.data() + txt.size()};
  const JVal root = ps.val();
  if (!ps.ok || root.k != JVal::OBJ) return set_error(WPIC_EINVAL, "cannot parse " + io->dir + name + ".json");
  FILE* f = std::fopen((io->dir + name + ".raw").c_str(), "rb");
  if (!f) return set_error(WPIC_EINVAL, "cannot open " + io->dir + name + ".raw");
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  const JVal* attr = root.get("attribute");
  long long off;
  std::vector<long long> shape;
  auto get_i4 = [&](const char* nm, int& v) {
    return meta_of(attr, nm, off, shape) && read_at(f, off, &v, 4);
  };
  auto get_r8 = [&](const char* nm, double* v, int n) {
    return meta_of(attr, nm, off, shape) && read_at(f, off, v, 8 * (size_t)n);
  };
  int indim, inp, inxgs, inxge, inygs, inyge, inzgs = io->nzgs, inzge = io->nzge, insp, inproc;
  bool ok = get_i4("it", *it) && get_i4("nxs", *nxs) && get_i4("nxe", *nxe) && get_i4("ndim", indim) &&
            get_i4("np", inp) && get_i4("nxgs", inxgs) && get_i4("nxge", inxge) && get_i4("nygs", inygs) &&
            get_i4("nyge", inyge);
  if (ok && io->dims == 3) ok = get_i4("nzgs", inzgs) && get_i4("nzge", inzge);
  ok = ok && get_i4("nsp", insp) && get_i4("nproc", inproc);
  if (!ok) return set_error(WPIC_EINVAL, "cannot read the attributes of " + name);
  // the reference reads delx, delt, c, r, q into its module state
  // (paraio.f90:311-321) before checking the grid; so does this
  if (insp != io->nsp) return set_error(WPIC_EINVAL, "*** Error: invalid input parameters ***");
  std::vector<double> rr(io->nsp), qq(io->nsp);
  double delx, delt, c;
  if (!(get_r8("delx", &delx, 1) && get_r8("delt", &delt, 1) && get_r8("c", &c, 1) && get_r8("r", rr.data(), io->nsp) &&
        get_r8("q", qq.data(), io->nsp)))
    return set_error(WPIC_EINVAL, "cannot read the attributes of " + name);
  io->delx = delx;
  io->delt = delt;
  io->c = c;
  io->r = rr;
  io->q = qq;
  if (!(inxgs == io->nxgs && inxge == io->nxge && inygs == io->nygs && inyge == io->nyge && inzgs == io->nzgs &&
        inzge == io->nzge && inp == io->np && inproc == io->nproc))
    return set_error(WPIC_EINVAL, "*** Error: invalid input parameters ***");
  const JVal* ds = root.get("dataset");
  const int nx = nx_of(io), ny = ny_of(io);
  const char* bad = "Fatal error in reading particle data";
  if (io->dims == 2) {
    if (!meta_of(ds, "up", off, shape) || shape != std::vector<long long>{io->ndim, io->np, ny, io->nsp, io->nproc})
      return set_error(WPIC_EINVAL, bad);
    if (!read_at(f, off, up, (size_t)8 * io->ndim * io->np * ny * io->nsp)) return set_error(WPIC_EINVAL, bad);
    if (!meta_of(ds, "np2", off, shape) || shape != std::vector<long long>{ny, io->nsp, io->nproc})
      return set_error(WPIC_EINVAL, bad);
    if (!read_at(f, off, np2, (size_t)4 * ny * io->nsp)) return set_error(WPIC_EINVAL, bad);
    if (!meta_of(ds, "uf", off, shape) || shape != std::vector<long long>{6, nx + 4, ny + 4, io->nproc})
      return set_error(WPIC_EINVAL, bad);
    if (!read_at(f, off, uf, (size_t)8 * 6 * (nx + 4) * (ny + 4))) return set_error(WPIC_EINVAL, bad);
    return 0;
  }
  const int nz = nz_of(io), ncol = ny * nz;
  if (!meta_of(ds, "np2", off, shape) || shape != std::vector<long long>{ny, nz, io->nsp, io->nproc})
    return set_error(WPIC_EINVAL, bad);
  if (!read_at(f, off, np2, (size_t)4 * ncol * io->nsp)) return set_error(WPIC_EINVAL, bad);
  // compact particles back into the padded columns in storage order
  std::memset(up, 0, sizeof(double) * (size_t)io->ndim * io->np * ncol * io->nsp);
  for (int s = 0; s < io->nsp; s++) {
    const std::string nm = "up" + two(s + 1);
    long long total = 0;
    for (int col = 0; col < ncol; col++) {
      const int n = np2[(size_t)col + (size_t)ncol * s];
      if (n < 0 || n > io->np) return set_error(WPIC_EMEM, "memory over (np2 > np)");
      total += n;
    }
    if (!meta_of(ds, nm.c_str(), off, shape) || shape.size() != 2 || shape[0] != io->ndim || shape[1] != total)
      return set_error(WPIC_EINVAL, bad);
    long long pos = off;
    for (int col = 0; col < ncol; col++) {
      const int n = np2[(size_t)col + (size_t)ncol * s];
      double* dst = up + (size_t)io->ndim * (size_t)io->np * ((size_t)col + (size_t)ncol * s);
      if (!read_at(f, pos, dst, (size_t)8 * io->ndim * n)) return set_error(WPIC_EINVAL, bad);
      pos += 8LL * io->ndim * n;
    }
  }
  if (!meta_of(ds, "uf", off, shape) || shape != std::vector<long long>{6, nx + 4, ny + 4, nz + 4, io->nproc})
    return set_error(WPIC_EINVAL, bad);
  if (!read_at(f, off, uf, (size_t)8 * 6 * (nx + 4) * (ny + 4) * (nz + 4))) return set_error(WPIC_EINVAL, bad);
  return 0;
}

// paraio__param: 2d/common/paraio.f90:431-550, 3d/common/paraio.f90:532-655
namespace {
int write_param(wpic_paraio* io, int n0, double wpe, double wpi, double wge, double wgi, double vti, double vte,
                const double* u0, const char* filename);
}

int wpic_paraio_param(wpic_paraio* io, int n0, double wpe, double wpi, double wge, double wgi, double vti, double vte,
                      const char* filename) {
  return write_param(io, n0, wpe, wpi, wge, wgi, vti, vte, nullptr, filename);
}

// the shock app's save_param (2d/proj/shock/app.f90:567-612): paraio__param's
// pair with the attribute u0 appended (JSON entry last, 8 bytes at the end of
// the raw file)
int wpic_paraio_param_shock(wpic_paraio* io, int n0, double wpe, double wpi, double wge, double wgi, double vti,
                            double vte, double u0, const char* filename) {
  return write_param(io, n0, wpe, wpi, wge, wgi, vti, vte, &u0, filename);
}

namespace {
int write_param(wpic_paraio* io, int n0, double wpe, double wpi, double wge, double wgi, double vti, double vte,
                const double* u0, const char* filename) {
  if (!io) return set_error(WPIC_EINVAL, "Initialize first by calling paraio__init()");
  if (!filename) return set_error(WPIC_EINVAL, "null argument");
  const std::string name = filename;
  const double vai = io->c * wgi / wpi;
  const double vae = io->c * wge / wpe;
  Writer w;
  int rc = open_writer(w, io, name);
  if (rc) return rc;
  w.begin_attributes();
  w.attr_i4("dummy_attribute", MOK);
  w.attr_i4("nx", nx_of(io));
  w.attr_i4("ny", ny_of(io));
  if (io->dims == 3) w.attr_i4("nz", nz_of(io));
  w.attr_i4("np", io->np);
  w.attr_r8("delx", io->delx);
  w.attr_r8("delt", io->delt);
  w.attr_r8("c", io->c);
  w.attr_r8v("r", io->r.data(), io->nsp);
  w.attr_r8v("q", io->q.data(), io->nsp);
  w.attr_r8("wpe", wpe);
  w.attr_r8("wge", wge);
  w.attr_r8("wpi", wpi);
  w.attr_r8("wgi", wgi);
  w.attr_r8("vti", vti);
  w.attr_r8("vte", vte);
  w.attr_r8("vai", vai);
  w.attr_r8("vae", vae);
  w.attr_i4("n0", n0);
  if (u0) w.attr_r8("u0", *u0);
  return close_writer(w, io, name);
}
}  // namespace

// paraio__mom: 2d/common/paraio.f90:555-713, 3d/common/paraio.f90:660-838.
// Divides mom(2:4) and mom(5:7) by mom(1) in place, as the reference does.
int wpic_paraio_mom(wpic_paraio* io, double* mom, const double* uf, int it) {
  if (!io) return set_error(WPIC_EINVAL, "Initialize first by calling paraio__init()");
  if (!mom || !uf) return set_error(WPIC_EINVAL, "null argument");
  const std::string name = it7(it) + "_mom";
  Writer w;
  int rc = open_writer(w, io, name);
  if (rc) return rc;
  w.begin_attributes();
  w.attr_i4("dummy_attribute", MOK);
  w.attr_i4("it", it);
  put_common(w, io);
  w.begin_datasets();
  const int nx = nx_of(io), ny = ny_of(io), nsp = io->nsp;
  const int nz = io->dims == 3 ? nz_of(io) : 1;
  const size_t nnode = io->dims == 3 ? (size_t)(nx + 2) * (ny + 2) * (nz + 2) : (size_t)(nx + 2) * (ny + 2);
  // mom(k,:,:,:) = mom(k,:,:,:) / mom(1,:,:,:) over the whole array, ghosts included
  for (size_t n = 0; n < nnode * nsp; n++)
    for (int k = 1; k < 7; k++) mom[7 * n + k] = mom[7 * n + k] / mom[7 * n];
  const size_t ncell = (size_t)nx * ny * nz;
  std::vector<double> den(ncell * nsp), vel(3 * ncell * nsp), temp(3 * ncell * nsp), f(6 * ncell);
  for (int s = 1; s <= nsp; s++)
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          const size_t cell = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k) + ncell * (s - 1);
          const double* m = io->dims == 3 ? &mom[i_mom3(io, 1, io->nxgs + i, io->nys + j, io->nzs + k, s)]
                                          : &mom[i_mom2(io, 1, io->nxgs + i, io->nys + j, s)];
          den[cell] = m[0];
          for (int d = 0; d < 3; d++) {
            vel[3 * cell + d] = m[1 + d];
            temp[3 * cell + d] = m[4 + d];
          }
        }
  if (io->dims == 2) {
    // cell-centred fields (paraio.f90:670-681), including the reference's
    // Bz average of uf(3,i+1,j) twice and without uf(3,i,j+1)
    for (int j = io->nys; j <= io->nye; j++)
      for (int i = io->nxgs; i <= io->nxge; i++) {
        double* t = &f[6 * ((size_t)(i - io->nxgs) + (size_t)nx * (j - io->nys))];
        auto U = [&](int c, int ii, int jj) { return uf[i_uf2(io, c, ii, jj)]; };
        t[0] = (U(1, i, j) + U(1, i, j + 1)) / 2;
        t[1] = (U(2, i, j) + U(2, i + 1, j)) / 2;
        t[2] = (U(3, i, j) + U(3, i + 1, j) + U(3, i + 1, j) + U(3, i + 1, j + 1)) / 4;
        t[3] = (U(4, i, j) + U(4, i + 1, j)) / 2;
        t[4] = (U(5, i, j) + U(5, i, j + 1)) / 2;
        t[5] = U(6, i, j);
      }
    w.dataset("den", "f8", {nx, ny, nsp}, den.data(), 8, "density");
    w.dataset("vel", "f8", {3, nx, ny, nsp}, vel.data(), 8, "velocity");
    w.dataset("temp", "f8", {3, nx, ny, nsp}, temp.data(), 8, "temperature");
    w.dataset("uf", "f8", {6, nx, ny}, f.data(), 8, "electromagnetic field");
  } else {
    // cell-centred fields (3d/common/paraio.f90:796-810)
    for (int k = io->nzs; k <= io->nze; k++)
      for (int j = io->nys; j <= io->nye; j++)
        for (int i = io->nxgs; i <= io->nxge; i++) {
          double* t = &f[6 * ((size_t)(i - io->nxgs) + (size_t)nx * ((size_t)(j - io->nys) + (size_t)ny * (k - io->nzs)))];
          auto U = [&](int c, int ii, int jj, int kk) { return uf[i_uf3(io, c, ii, jj, kk)]; };
          t[0] = (U(1, i, j, k) + U(1, i, j + 1, k) + U(1, i, j, k + 1) + U(1, i, j + 1, k + 1)) / 4;
          t[1] = (U(2, i, j, k) + U(2, i + 1, j, k) + U(2, i, j, k + 1) + U(2, i + 1, j, k + 1)) / 4;
          t[2] = (U(3, i, j, k) + U(3, i + 1, j, k) + U(3, i, j + 1, k) + U(3, i + 1, j + 1, k)) / 4;
          t[3] = (U(4, i, j, k) + U(4, i + 1, j, k)) / 2;
          t[4] = (U(5, i, j, k) + U(5, i, j + 1, k)) / 2;
          t[5] = (U(6, i, j, k) + U(6, i, j, k + 1)) / 2;
        }
    w.dataset("den", "f8", {nx, ny, nz, nsp}, den.data(), 8, "density");
    w.dataset("vel", "f8", {3, nx, ny, nz, nsp}, vel.data(), 8, "velocity");
    w.dataset("temp", "f8", {3, nx, ny, nz, nsp}, temp.data(), 8, "temperature");
    w.dataset("uf", "f8", {6, nx, ny, nz}, f.data(), 8, "electromagnetic field");
  }
  return close_writer(w, io, name);
}

// paraio__ptcl / paraio__orb: 2d/common/paraio.f90:718-741 via write_particle
// :746-893; 3d/common/paraio.f90:843-1002
int wpic_paraio_ptcl(wpic_paraio* io, const double* up, const double* uf, const int* np2, int it) {
  if (!io) return set_error(WPIC_EINVAL, "Initialize first by calling paraio__init()");
  if (!up || !uf || !np2) return set_error(WPIC_EINVAL, "null argument");
  return write_particles(io, up, uf, np2, it, 0, "_ptcl");
}

int wpic_paraio_orb(wpic_paraio* io, const double* up, const double* uf, const int* np2, int it) {
  if (!io) return set_error(WPIC_EINVAL, "Initialize first by calling paraio__init()");
  if (!up || !uf || !np2) return set_error(WPIC_EINVAL, "null argument");
  return write_particles(io, up, uf, np2, it, 1, "_orb");
}

}  // extern "C"

// vacc_vlsv.cpp -- VLSV container and the velocity block arrays of one population
// (include/vacc_vlsv.h).
//
// Host C++ only.  The block values move through the streamed transfer of vacc_blocks_io.h:
// while one pinned chunk is written to (read from) the file, the next one is in flight
// between HBM and host memory.
#include <errno.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/vacc_vlsv.h"
#include "vacc_blocks_io.h"

using vacc::set_error;

namespace {

constexpr uint64_t HEADER_BYTES = 16;
const char* const SPATIAL_MESH = "SpatialGrid";   // iowrite.cpp:149
const char* const VMESH_TYPE = "amr_ucd";         // vlsv::mesh::STRING_UCD_AMR, iowrite.cpp:187

struct Entry {
   std::string tag;
   std::map<std::string, std::string> attr;   // every attribute, the mandatory ones included
   uint64_t arraysize = 0, vectorsize = 0, datasize = 0, offset = 0;
   int32_t datatype = -1;                      // -1: not an array tag
   uint64_t bytes() const { return arraysize * vectorsize * datasize; }
};

using Attrs = std::vector<std::pair<std::string, std::string>>;

const char* type_name(int32_t t) { return t == VACC_VLSV_INT ? "int" : (t == VACC_VLSV_UINT ? "uint" : "float"); }

int32_t type_of(const std::string& s) {
   if (s == "int") return VACC_VLSV_INT;
   if (s == "uint") return VACC_VLSV_UINT;
   if (s == "float") return VACC_VLSV_FLOAT;
   return -1;
}

bool valid_size(int32_t t, uint64_t ds) {
   if (t == VACC_VLSV_FLOAT) return ds == 4 || ds == 8;
   if (t == VACC_VLSV_INT || t == VACC_VLSV_UINT) return ds == 1 || ds == 2 || ds == 4 || ds == 8;
   return false;
}

bool is_mandatory(const std::string& k) {
   return k == "arraysize" || k == "vectorsize" || k == "datatype" || k == "datasize";
}

bool clean(const std::string& s) { return s.find_first_of("<>&\"") == std::string::npos; }

int io_error(const char* what) { return set_error(VACC_E_ARG, std::string(what) + ": " + strerror(errno)); }

// Positioned I/O of n bytes at off; returns 0, an errno value, or -1 at an early end of file.
int pio(int fd, char* b, size_t n, uint64_t off, bool wr) {
   while (n) {
      const ssize_t k = wr ? ::pwrite(fd, b, n, (off_t)off) : ::pread(fd, b, n, (off_t)off);
      if (k < 0) {
         if (errno == EINTR) continue;
         return errno;
      }
      if (k == 0) return -1;
      b += k;
      n -= (size_t)k;
      off += (uint64_t)k;
   }
   return 0;
}

// Large transfers are split over IO_THREADS concurrent calls: a page-cache copy on one thread
// runs at a fraction of the host's memory bandwidth.
constexpr int IO_THREADS = 4;
constexpr size_t IO_SPLIT_MIN = size_t(8) << 20;

int pio_all(int fd, void* p, size_t n, uint64_t off, bool wr) {
   char* b = static_cast<char*>(p);
   int err = 0;
   if (n < IO_SPLIT_MIN) {
      err = pio(fd, b, n, off, wr);
   } else {
      const size_t part = ((n + IO_THREADS - 1) / IO_THREADS + 4095) & ~size_t(4095);
      int rcs[IO_THREADS] = {0};
      std::thread ts[IO_THREADS];
      for (int t = 1; t < IO_THREADS; ++t) {
         const size_t a = std::min(n, t * part), e = std::min(n, (t + 1) * part);
         if (a < e) ts[t] = std::thread([=, &rcs] { rcs[t] = pio(fd, b + a, e - a, off + a, wr); });
      }
      rcs[0] = pio(fd, b, std::min(n, part), off, wr);
      for (int t = 1; t < IO_THREADS; ++t)
         if (ts[t].joinable()) ts[t].join();
      for (int t = 0; t < IO_THREADS && !err; ++t) err = rcs[t];
   }
   if (err == -1) return set_error(VACC_E_ARG, "unexpected end of file");
   if (err) return set_error(VACC_E_ARG, std::string(wr ? "write failed: " : "read failed: ") + strerror(err));
   return 0;
}

int pread_all(int fd, void* p, size_t n, uint64_t off) { return pio_all(fd, p, n, off, false); }
int pwrite_all(int fd, const void* p, size_t n, uint64_t off) { return pio_all(fd, const_cast<void*>(p), n, off, true); }

int attrs_from(uint32_t nattr, const char* const* attr, Attrs& out) {
   if (nattr && !attr) return set_error(VACC_E_ARG, "null attribute list");
   for (uint32_t i = 0; i < nattr; ++i) {
      if (!attr[2 * i] || !attr[2 * i + 1]) return set_error(VACC_E_ARG, "null attribute string");
      out.emplace_back(attr[2 * i], attr[2 * i + 1]);
   }
   return 0;
}

// Element conversion of one stored value (datatype, datasize) at p.
template <typename T>
T load_as(const char* p, int32_t t, uint64_t ds) {
   switch (t) {
      case VACC_VLSV_FLOAT:
         if (ds == 4) { float v; memcpy(&v, p, 4); return (T)v; }
         { double v; memcpy(&v, p, 8); return (T)v; }
      case VACC_VLSV_UINT:
         switch (ds) {
            case 1: return (T)(uint8_t)p[0];
            case 2: { uint16_t v; memcpy(&v, p, 2); return (T)v; }
            case 4: { uint32_t v; memcpy(&v, p, 4); return (T)v; }
            default: { uint64_t v; memcpy(&v, p, 8); return (T)v; }
         }
      default:
         switch (ds) {
            case 1: return (T)(int8_t)p[0];
            case 2: { int16_t v; memcpy(&v, p, 2); return (T)v; }
            case 4: { int32_t v; memcpy(&v, p, 4); return (T)v; }
            default: { int64_t v; memcpy(&v, p, 8); return (T)v; }
         }
   }
}

}  // namespace

struct vacc_vlsv_writer {
   int fd = -1;
   uint64_t offset = HEADER_BYTES;
   std::vector<Entry> entries;
};

struct vacc_vlsv_reader {
   int fd = -1;
   uint64_t size = 0;
   std::vector<Entry> entries;
};

namespace {

int make_entry(const char* tag, const Attrs& attrs, uint64_t arraysize, uint64_t vectorsize, int32_t datatype,
               uint64_t datasize, Entry& e) {
   if (!tag || !*tag || !clean(tag) || strpbrk(tag, " \t\n/=")) return set_error(VACC_E_ARG, "bad tag name");
   if (!valid_size(datatype, datasize)) return set_error(VACC_E_ARG, "unsupported datatype/datasize");
   e.tag = tag;
   for (const auto& kv : attrs) {
      if (is_mandatory(kv.first)) return set_error(VACC_E_ARG, "attribute " + kv.first + " is set by the writer");
      if (kv.first.empty() || !clean(kv.first) || kv.first.find_first_of(" \t\n/=") != std::string::npos ||
          !clean(kv.second))
         return set_error(VACC_E_ARG, "bad attribute " + kv.first);
      e.attr[kv.first] = kv.second;
   }
   e.arraysize = arraysize;
   e.vectorsize = vectorsize;
   e.datatype = datatype;
   e.datasize = datasize;
   e.attr["arraysize"] = std::to_string(arraysize);
   e.attr["vectorsize"] = std::to_string(vectorsize);
   e.attr["datatype"] = type_name(datatype);
   e.attr["datasize"] = std::to_string(datasize);
   return 0;
}

int put_array(vacc_vlsv_writer* w, const char* tag, const Attrs& attrs, uint64_t arraysize, uint64_t vectorsize,
              int32_t datatype, uint64_t datasize, const void* data) {
   Entry e;
   int rc = make_entry(tag, attrs, arraysize, vectorsize, datatype, datasize, e);
   if (rc) return rc;
   if (e.bytes() && !data) return set_error(VACC_E_ARG, "null array data");
   e.offset = w->offset;
   if ((rc = pwrite_all(w->fd, data, e.bytes(), w->offset))) return rc;
   w->offset += e.bytes();
   w->entries.push_back(std::move(e));
   return 0;
}

std::string footer_text(const std::vector<Entry>& entries) {
   std::string x = "<VLSV>\n";
   for (const auto& e : entries) {
      x += "   <" + e.tag;
      for (const auto& kv : e.attr) x += " " + kv.first + "=\"" + kv.second + "\"";
      x += ">" + std::to_string(e.offset) + "</" + e.tag + ">\n";
   }
   x += "</VLSV>\n";
   return x;
}

std::string unescape(const std::string& s) {
   static const std::pair<const char*, char> ents[] = {{"&quot;", '"'}, {"&apos;", '\''}, {"&lt;", '<'},
                                                      {"&gt;", '>'}, {"&amp;", '&'}};
   std::string out;
   for (size_t i = 0; i < s.size();) {
      bool hit = false;
      if (s[i] == '&')
         for (const auto& en : ents) {
            const size_t n = strlen(en.first);
            if (s.compare(i, n, en.first) == 0) {
               out += en.second;
               i += n;
               hit = true;
               break;
            }
         }
      if (!hit) out += s[i++];
   }
   return out;
}

bool space(char ch) { return ch == ' ' || ch == '\t' || ch == '\n' || ch == '\r'; }

// Parses the children of the <VLSV> root: <TAG key="value" ...>text</TAG> or <TAG .../>.
bool parse_footer(const std::string& x, std::vector<Entry>& out) {
   size_t p = x.find("<VLSV>");
   if (p == std::string::npos) return false;
   p += 6;
   for (;;) {
      p = x.find('<', p);
      if (p == std::string::npos) return false;
      if (x.compare(p, 7, "</VLSV>") == 0) return true;
      if (p + 1 < x.size() && (x[p + 1] == '?' || x[p + 1] == '!')) {
         p = x.find('>', p);
         if (p == std::string::npos) return false;
         continue;
      }
      size_t q = p + 1;
      while (q < x.size() && !space(x[q]) && x[q] != '>' && x[q] != '/') ++q;
      if (q >= x.size()) return false;
      Entry e;
      e.tag = x.substr(p + 1, q - p - 1);
      bool closed = false;
      for (;;) {
         while (q < x.size() && space(x[q])) ++q;
         if (q >= x.size()) return false;
         if (x[q] == '>') {
            ++q;
            break;
         }
         if (x[q] == '/' && q + 1 < x.size() && x[q + 1] == '>') {
            q += 2;
            closed = true;
            break;
         }
         const size_t k0 = q;
         while (q < x.size() && x[q] != '=' && !space(x[q])) ++q;
         const std::string key = x.substr(k0, q - k0);
         while (q < x.size() && space(x[q])) ++q;
         if (q >= x.size() || x[q] != '=') return false;
         ++q;
         while (q < x.size() && space(x[q])) ++q;
         if (q >= x.size() || (x[q] != '"' && x[q] != '\'')) return false;
         const char quote = x[q++];
         const size_t v1 = x.find(quote, q);
         if (v1 == std::string::npos) return false;
         e.attr[key] = unescape(x.substr(q, v1 - q));
         q = v1 + 1;
      }
      std::string text;
      if (!closed) {
         const size_t end = x.find("</" + e.tag, q);
         if (end == std::string::npos) return false;
         text = x.substr(q, end - q);
         q = x.find('>', end);
         if (q == std::string::npos) return false;
         ++q;
      }
      p = q;
      auto num = [&](const char* k, uint64_t& v) {
         auto it = e.attr.find(k);
         if (it == e.attr.end()) return false;
         char* endp = nullptr;
         v = strtoull(it->second.c_str(), &endp, 10);
         return endp && *endp == '\0' && !it->second.empty();
      };
      auto it = e.attr.find("datatype");
      char* endp = nullptr;
      const size_t t0 = text.find_first_not_of(" \t\r\n");
      const std::string t = t0 == std::string::npos ? "" : text.substr(t0, text.find_last_not_of(" \t\r\n") - t0 + 1);
      const uint64_t off = strtoull(t.c_str(), &endp, 10);
      if (it != e.attr.end() && num("arraysize", e.arraysize) && num("vectorsize", e.vectorsize) &&
          num("datasize", e.datasize) && !t.empty() && endp && *endp == '\0') {
         e.datatype = type_of(it->second);
         e.offset = off;
         if (!valid_size(e.datatype, e.datasize)) e.datatype = -1;
      }
      out.push_back(std::move(e));
   }
}

const Entry* find(const vacc_vlsv_reader* r, const char* tag, const Attrs& attrs) {
   for (const auto& e : r->entries) {
      if (e.datatype < 0 || e.tag != tag) continue;
      bool ok = true;
      for (const auto& kv : attrs) {
         auto it = e.attr.find(kv.first);
         if (it == e.attr.end() || it->second != kv.second) {
            ok = false;
            break;
         }
      }
      if (ok) return &e;
   }
   return nullptr;
}

// The array's byte size, or false when arraysize * vectorsize * datasize overflows.
bool array_bytes(const Entry& e, uint64_t& n) {
   uint64_t v;
   return !__builtin_mul_overflow(e.arraysize, e.vectorsize, &v) && !__builtin_mul_overflow(v, e.datasize, &n);
}

int need(const vacc_vlsv_reader* r, const char* tag, const Attrs& attrs, const Entry*& e) {
   e = find(r, tag, attrs);
   if (!e) {
      std::string a;
      for (const auto& kv : attrs) a += " " + kv.first + "=" + kv.second;
      return set_error(VACC_E_ARG, std::string("no array ") + tag + a + " in the file");
   }
   uint64_t n = 0;
   if (!array_bytes(*e, n) || e->offset > r->size || n > r->size - e->offset)
      return set_error(VACC_E_ARG, std::string("array ") + tag + " beyond end of file");
   return 0;
}

int read_raw(const vacc_vlsv_reader* r, const Entry& e, uint64_t begin, uint64_t amount, void* out) {
   if (begin > e.arraysize || amount > e.arraysize - begin)
      return set_error(VACC_E_ARG, "read of " + e.tag + " beyond its arraysize");
   const uint64_t vb = e.vectorsize * e.datasize;
   return pread_all(r->fd, out, amount * vb, e.offset + begin * vb);
}

// Vectors [begin, begin + amount) of e converted to T, vectorsize components each.
template <typename T>
int read_as(const vacc_vlsv_reader* r, const Entry& e, uint64_t begin, uint64_t amount, std::vector<T>& out) {
   std::vector<char> raw(amount * e.vectorsize * e.datasize);
   int rc = read_raw(r, e, begin, amount, raw.data());
   if (rc) return rc;
   out.resize(amount * e.vectorsize);
   for (size_t i = 0; i < out.size(); ++i) out[i] = load_as<T>(raw.data() + i * e.datasize, e.datatype, e.datasize);
   return 0;
}

}  // namespace

extern "C" {

int vacc_vlsv_create(const char* path, vacc_vlsv_writer** out) {
   if (!path || !out) return set_error(VACC_E_ARG, "null argument");
   const int fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
   if (fd < 0) return io_error("cannot create file");
   const uint64_t head[2] = {0, 0};   // endianness word (byte 0 = 0: little-endian), footer offset
   int rc = pwrite_all(fd, head, sizeof head, 0);
   if (rc) {
      ::close(fd);
      return rc;
   }
   auto* w = new vacc_vlsv_writer();
   w->fd = fd;
   *out = w;
   return 0;
}

int vacc_vlsv_write_array(vacc_vlsv_writer* w, const char* tag, uint32_t nattr, const char* const* attr,
                          uint64_t arraysize, uint64_t vectorsize, int32_t datatype, uint64_t datasize,
                          const void* data) {
   if (!w) return set_error(VACC_E_ARG, "null writer");
   Attrs a;
   int rc = attrs_from(nattr, attr, a);
   if (rc) return rc;
   return put_array(w, tag, a, arraysize, vectorsize, datatype, datasize, data);
}

// iowrite.cpp:141-318, one population, one process.
int vacc_vlsv_write_velocity_distribution(vacc_vlsv_writer* w, vacc_ctx* ctx, const char* pop_name,
                                          const uint64_t* cell_ids) {
   if (!w || !ctx || !pop_name) return set_error(VACC_E_ARG, "null argument");
   if (!*pop_name || !clean(pop_name)) return set_error(VACC_E_ARG, "bad population name");
   std::vector<uint32_t> counts, gids;
   int rc = vacc::export_lists(ctx, counts, gids);
   if (rc) return rc;
   const uint64_t ncells = counts.size(), nblocks = gids.size();
   std::vector<uint64_t> ids(ncells);
   for (uint64_t i = 0; i < ncells; ++i) ids[i] = cell_ids ? cell_ids[i] : i + 1;
   const Attrs cell_attrs = {{"mesh", SPATIAL_MESH}, {"name", pop_name}};
   if ((rc = put_array(w, "CELLSWITHBLOCKS", cell_attrs, ncells, 1, VACC_VLSV_UINT, 8, ids.data()))) return rc;
   if ((rc = put_array(w, "BLOCKSPERCELL", cell_attrs, ncells, 1, VACC_VLSV_UINT, 4, counts.data()))) return rc;
   // velocity mesh: bounding box in blocks and cells per block, node coordinates per axis
   vacc_mesh m;
   double cell[3];
   vacc::ctx_mesh(ctx, &m, cell);
   const uint64_t bbox[6] = {m.grid_length[0], m.grid_length[1], m.grid_length[2], VACC_WID, VACC_WID, VACC_WID};
   const Attrs mesh_attrs = {{"mesh", pop_name}, {"type", VMESH_TYPE}, {"max_velocity_ref_level", "0"}};
   if ((rc = put_array(w, "MESH_BBOX", mesh_attrs, 6, 1, VACC_VLSV_UINT, 8, bbox))) return rc;
   static const char* const crd_tags[3] = {"MESH_NODE_CRDS_X", "MESH_NODE_CRDS_Y", "MESH_NODE_CRDS_Z"};
   for (int d = 0; d < 3; ++d) {
      const uint64_t nn = (uint64_t)m.grid_length[d] * VACC_WID + 1;
      std::vector<double> crds(nn);
      for (uint64_t i = 0; i < nn; ++i) crds[i] = m.min_limits[d] + (double)i * cell[d];
      if ((rc = put_array(w, crd_tags[d], mesh_attrs, nn, 1, VACC_VLSV_FLOAT, 8, crds.data()))) return rc;
   }
   if ((rc = put_array(w, "BLOCKIDS", cell_attrs, nblocks, 1, VACC_VLSV_UINT, 4, gids.data()))) return rc;
   // f, streamed from HBM in chunks
   Entry e;
   if ((rc = make_entry("BLOCKVARIABLE", cell_attrs, nblocks, VACC_WID3, VACC_VLSV_FLOAT, 4, e))) return rc;
   e.offset = w->offset;
   rc = vacc::export_values(ctx, [&](const float* f, uint64_t first, uint64_t n) {
      return pwrite_all(w->fd, f, n * VACC_WID3 * sizeof(float), e.offset + first * VACC_WID3 * sizeof(float));
   });
   if (rc) return rc;
   w->offset += e.bytes();
   w->entries.push_back(std::move(e));
   return 0;
}

int vacc_vlsv_close(vacc_vlsv_writer* w) {
   if (!w) return set_error(VACC_E_ARG, "null writer");
   const std::string x = footer_text(w->entries);
   int rc = pwrite_all(w->fd, x.data(), x.size(), w->offset);
   if (!rc) rc = pwrite_all(w->fd, &w->offset, sizeof(uint64_t), 8);
   if (::close(w->fd) != 0 && !rc) rc = io_error("close failed");
   delete w;
   return rc;
}

int vacc_vlsv_open(const char* path, vacc_vlsv_reader** out) {
   if (!path || !out) return set_error(VACC_E_ARG, "null argument");
   const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
   if (fd < 0) return io_error("cannot open file");
   auto* r = new vacc_vlsv_reader();
   r->fd = fd;
   struct stat st;
   uint64_t head[2] = {0, 0};
   int rc = 0;
   if (fstat(fd, &st) != 0) rc = io_error("stat failed");
   r->size = (uint64_t)st.st_size;
   if (!rc && r->size < HEADER_BYTES) rc = set_error(VACC_E_ARG, "not a VLSV file (too short)");
   if (!rc) rc = pread_all(fd, head, sizeof head, 0);
   if (!rc) {
      // footer offset at byte 8 (endianness word first), else the older layout at byte 0
      bool ok = false;
      for (uint64_t off : {head[1], head[0]}) {
         if (off < 8 || off >= r->size) continue;
         std::string x(r->size - off, '\0');
         if (pread_all(fd, &x[0], x.size(), off)) continue;
         const size_t s0 = x.find_first_not_of(" \t\r\n");
         if (s0 == std::string::npos || x.compare(s0, 6, "<VLSV>") != 0) continue;
         r->entries.clear();
         if (parse_footer(x, r->entries)) {
            ok = true;
            break;
         }
      }
      if (!ok) rc = set_error(VACC_E_ARG, "not a VLSV file (no footer)");
   }
   if (rc) {
      ::close(fd);
      delete r;
      return rc;
   }
   *out = r;
   return 0;
}

int vacc_vlsv_array_info(vacc_vlsv_reader* r, const char* tag, uint32_t nattr, const char* const* attr,
                         uint64_t* arraysize, uint64_t* vectorsize, int32_t* datatype, uint64_t* datasize) {
   if (!r || !tag) return set_error(VACC_E_ARG, "null argument");
   Attrs a;
   int rc = attrs_from(nattr, attr, a);
   const Entry* e = nullptr;
   if (rc || (rc = need(r, tag, a, e))) return rc;
   if (arraysize) *arraysize = e->arraysize;
   if (vectorsize) *vectorsize = e->vectorsize;
   if (datatype) *datatype = e->datatype;
   if (datasize) *datasize = e->datasize;
   return 0;
}

int vacc_vlsv_read_array(vacc_vlsv_reader* r, const char* tag, uint32_t nattr, const char* const* attr,
                         uint64_t begin, uint64_t amount, void* out) {
   if (!r || !tag || (amount && !out)) return set_error(VACC_E_ARG, "null argument");
   Attrs a;
   int rc = attrs_from(nattr, attr, a);
   const Entry* e = nullptr;
   if (rc || (rc = need(r, tag, a, e))) return rc;
   return read_raw(r, *e, begin, amount, out);
}

// ioread.cpp:488-700 (with _readBlockData, ioread.cpp:371-486), one population, one process.
int vacc_vlsv_read_block_data(vacc_vlsv_reader* r, vacc_ctx* ctx, const char* pop_name, const char* mesh_name,
                              uint64_t cell_offset, uint32_t ncells, uint64_t* cell_ids_out) {
   if (!r || !ctx || !pop_name || !mesh_name) return set_error(VACC_E_ARG, "null argument");
   vacc_mesh m;
   double cell[3];
   vacc::ctx_mesh(ctx, &m, cell);
   const Entry* e = nullptr;
   int rc = need(r, "MESH_BBOX", {{"mesh", pop_name}}, e);
   if (rc) return rc;
   if (e->arraysize * e->vectorsize < 6 || e->arraysize * e->vectorsize > 64)
      return set_error(VACC_E_ARG, "MESH_BBOX does not hold 6 values");
   std::vector<uint64_t> bbox;
   if ((rc = read_as(r, *e, 0, e->arraysize, bbox))) return rc;
   uint64_t shift[3] = {0, 0, 0};
   const bool renumber = bbox[0] != m.grid_length[0] || bbox[1] != m.grid_length[1] || bbox[2] != m.grid_length[2];
   if (renumber) {
      // ioread.cpp:520-631: the file's velocity mesh must sit inside ours at a block offset
      if (bbox[3] != VACC_WID || bbox[4] != VACC_WID || bbox[5] != VACC_WID)
         return set_error(VACC_E_ARG, "file velocity blocks are not 4x4x4 cells");
      if (bbox[0] == 0 || bbox[1] == 0 || bbox[2] == 0) return set_error(VACC_E_ARG, "empty file velocity mesh");
      for (int d = 0; d < 3; ++d)
         if (m.grid_length[d] < bbox[d]) return set_error(VACC_E_ARG, "trying to shrink velocity space");
      static const char* const crd_tags[3] = {"MESH_NODE_CRDS_X", "MESH_NODE_CRDS_Y", "MESH_NODE_CRDS_Z"};
      for (int d = 0; d < 3; ++d) {
         const Entry* ce = nullptr;
         if ((rc = need(r, crd_tags[d], {{"mesh", pop_name}}, ce))) return rc;
         const uint64_t nn = bbox[d] * bbox[d + 3] + 1;
         if (ce->vectorsize != 1 || ce->arraysize < nn) return set_error(VACC_E_ARG, "short node coordinate array");
         std::vector<double> crds;
         if ((rc = read_as(r, *ce, 0, nn, crds))) return rc;
         for (double c : crds) {
            const double ci = (c - m.min_limits[d]) / cell[d];
            if (std::fabs(std::nearbyint(ci) - ci) > 1.0 / 10000.0)
               return set_error(VACC_E_ARG, "cannot resize velocity space: cell coordinates do not match");
         }
         const double c0 = std::nearbyint((crds[0] - m.min_limits[d]) / cell[d]);
         if (c0 < 0.0 || std::fmod(c0, (double)VACC_WID) != 0.0)
            return set_error(VACC_E_ARG, "the file's velocity space must start at a block boundary of ours");
         shift[d] = (uint64_t)c0 / VACC_WID;
         if (shift[d] + bbox[d] > m.grid_length[d])
            return set_error(VACC_E_ARG, "the file's velocity space does not fit into ours");
      }
   }
   const Attrs cell_attrs = {{"mesh", mesh_name}, {"name", pop_name}};
   if ((rc = need(r, "CELLSWITHBLOCKS", cell_attrs, e))) return rc;
   if (cell_offset > e->arraysize || ncells > e->arraysize - cell_offset)
      return set_error(VACC_E_ARG, "cell range beyond CELLSWITHBLOCKS");
   if (cell_ids_out && ncells) {
      std::vector<uint64_t> ids;
      if ((rc = read_as(r, *e, cell_offset, ncells, ids))) return rc;
      std::copy(ids.begin(), ids.end(), cell_ids_out);
   }
   const Entry* bpc = nullptr;
   if ((rc = need(r, "BLOCKSPERCELL", cell_attrs, bpc))) return rc;
   if (bpc->vectorsize != 1 || cell_offset + ncells > bpc->arraysize)
      return set_error(VACC_E_ARG, "cell range beyond BLOCKSPERCELL");
   std::vector<uint64_t> per;
   if ((rc = read_as(r, *bpc, 0, cell_offset + ncells, per))) return rc;
   uint64_t start = 0;
   for (uint64_t i = 0; i < cell_offset; ++i) start += per[i];
   std::vector<uint64_t> offsets(ncells + 1, 0);
   for (uint32_t i = 0; i < ncells; ++i) offsets[i + 1] = offsets[i] + per[cell_offset + i];
   const uint64_t nb = offsets[ncells];
   const Entry* ide = nullptr;
   if ((rc = need(r, "BLOCKIDS", cell_attrs, ide))) return rc;
   if (ide->datasize != sizeof(uint32_t) || ide->vectorsize != 1)
      return set_error(VACC_E_ARG, "BlockID data size does not match");   // ioread.cpp:422-425
   std::vector<uint32_t> gids;
   if ((rc = read_as(r, *ide, start, nb, gids))) return rc;
   if (renumber) {
      for (auto& g : gids) {   // ioread.cpp:617-629
         const uint64_t x = g % bbox[0], y = (g / bbox[0]) % bbox[1], z = g / (bbox[0] * bbox[1]);
         if (z >= bbox[2]) return set_error(VACC_E_ARG, "GID outside the file's velocity mesh");
         g = (uint32_t)((x + shift[0]) + (y + shift[1]) * m.grid_length[0] +
                        (z + shift[2]) * (uint64_t)m.grid_length[0] * m.grid_length[1]);
      }
   }
   {
      std::vector<uint32_t> t;
      for (uint32_t i = 0; i < ncells; ++i) {
         t.assign(gids.begin() + offsets[i], gids.begin() + offsets[i + 1]);
         std::sort(t.begin(), t.end());
         if (std::adjacent_find(t.begin(), t.end()) != t.end())
            return set_error(VACC_E_ARG, "duplicate block GID in a cell");
      }
   }
   const Entry* fe = nullptr;
   if ((rc = need(r, "BLOCKVARIABLE", cell_attrs, fe))) return rc;
   if (fe->vectorsize != VACC_WID3) return set_error(VACC_E_ARG, "Blocksize does not match in the file");
   if (fe->datasize != 4 && fe->datasize != 8) return set_error(VACC_E_ARG, "unsupported BLOCKVARIABLE data size");
   if (start > fe->arraysize || nb > fe->arraysize - start) return set_error(VACC_E_ARG, "BLOCKVARIABLE too short");
   const Entry fv = *fe;
   std::vector<char> raw;
   auto source = [&](float* f, uint64_t first, uint64_t n) -> int {
      if (fv.datatype == VACC_VLSV_FLOAT && fv.datasize == 4) return read_raw(r, fv, start + first, n, f);
      raw.resize(n * VACC_WID3 * fv.datasize);
      const int rc2 = read_raw(r, fv, start + first, n, raw.data());
      if (rc2) return rc2;
      for (uint64_t i = 0; i < n * VACC_WID3; ++i) f[i] = load_as<float>(raw.data() + i * fv.datasize, fv.datatype, fv.datasize);
      return 0;
   };
   return vacc::import_blocks(ctx, ncells, offsets.data(), gids.data(), source);
}

int vacc_vlsv_reader_close(vacc_vlsv_reader* r) {
   if (!r) return 0;
   ::close(r->fd);
   delete r;
   return 0;
}

}  // extern "C"

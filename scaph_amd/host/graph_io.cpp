// graph_io.cpp — readers/writers for the reference's graph files and the
// page cut rule.  Layouts and citations are in graph_io.h.
#include "graph_io.h"

#include <sys/stat.h>

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

namespace scaph {

namespace {

int64_t file_size(const std::string& path) {
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return -1;
  return (int64_t)st.st_size;
}

bool write_file(const std::string& path, const void* data, size_t bytes, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    if (err) *err = "cannot create " + path;
    return false;
  }
  const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
  const bool ok = w == bytes && std::fclose(f) == 0;
  if (!ok && err) *err = "short write to " + path;
  return ok;
}

bool read_file(const std::string& path, void* data, size_t bytes, std::string* err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (err) *err = "cannot open " + path;
    return false;
  }
  const size_t r = bytes ? std::fread(data, 1, bytes, f) : 0;
  std::fclose(f);
  if (r != bytes) {
    if (err) *err = "short read from " + path;
    return false;
  }
  return true;
}

bool fits(const uint64_t* off, uint32_t left, uint32_t right, uint64_t P, bool weighted) {
  const uint64_t edges = off[(uint64_t)right + 1] - off[left];
  return edges * (weighted ? 2 : 1) + 3ull * ((uint64_t)right - left + 1) < P;
}

}  // namespace

bool page_cut(const uint64_t* off, uint32_t n, uint32_t left, uint64_t P, bool weighted,
              uint32_t* right) {
  if (left >= n) return false;
  if (!fits(off, left, left, P, weighted)) return false;
  uint64_t hi = std::min<uint64_t>((uint64_t)left + P, (uint64_t)n - 1);
  if (fits(off, left, (uint32_t)hi, P, weighted)) {
    *right = (uint32_t)hi;
    return true;
  }
  // fits(lo) true, fits(hi) false: find the boundary
  uint64_t lo = left;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (fits(off, left, (uint32_t)mid, P, weighted)) lo = mid; else hi = mid;
  }
  *right = (uint32_t)lo;
  return true;
}

bool page_cuts_raw(const uint64_t* off, uint32_t n, uint64_t P, bool weighted,
                   std::vector<scaph_page>* pages, std::string* err) {
  pages->clear();
  uint32_t left = 0;
  while (left < n) {
    uint32_t right;
    if (!page_cut(off, n, left, P, weighted, &right)) {
      if (err)
        *err = "vertex " + std::to_string(left) + " (degree " +
               std::to_string(off[left + 1] - off[left]) + ") does not fit a page of " +
               std::to_string(P) + " words";
      return false;
    }
    scaph_page p;
    p.left = left;
    p.right = right;
    p.nodenum = right - left + 1;
    p.edgenum = (uint32_t)(off[(uint64_t)right + 1] - off[left]);
    pages->push_back(p);
    left = right + 1;
    if (right == UINT32_MAX) break;
  }
  return true;
}

bool page_cuts(const HostCsr& g, uint64_t P, bool weighted, std::vector<scaph_page>* pages,
               std::string* err) {
  return page_cuts_raw(g.offsets.data(), g.n, P, weighted, pages, err);
}

void fill_page(const HostCsr& g, const scaph_page& p, bool weighted, uint32_t* buf) {
  const uint64_t base = g.offsets[p.left];
  for (uint32_t i = 0; i < p.nodenum; ++i) {
    const uint32_t v = p.left + i;
    buf[3ull * i] = v;
    buf[3ull * i + 1] = (uint32_t)(g.offsets[v] - base);
    buf[3ull * i + 2] = (uint32_t)(g.offsets[(uint64_t)v + 1] - g.offsets[v]);
  }
  uint32_t* e = buf + 3ull * p.nodenum;
  if (weighted) {
    for (uint32_t k = 0; k < p.edgenum; ++k) {
      e[2ull * k] = g.ngh[base + k];
      e[2ull * k + 1] = g.wgh.empty() ? 0u : g.wgh[base + k];
    }
  } else if (p.edgenum) {
    std::memcpy(e, g.ngh.data() + base, (size_t)p.edgenum * 4);
  }
}

bool write_csr_files(const std::string& base, const HostCsr& g, bool idx64, bool write_wgh,
                     std::string* err) {
  const uint64_t E = g.num_edges();
  if (idx64) {
    if (!write_file(base + "_csr.idx", g.offsets.data(), g.offsets.size() * 8, err)) return false;
  } else {
    if (E > UINT32_MAX) {
      if (err) *err = "graph has more than 2^32 edges: u32 row offsets cannot hold it";
      return false;
    }
    std::vector<uint32_t> o32(g.offsets.begin(), g.offsets.end());
    if (!write_file(base + "_csr.idx", o32.data(), o32.size() * 4, err)) return false;
  }
  if (!write_file(base + "_csr.ngh", g.ngh.data(), E * 4, err)) return false;
  std::vector<uint32_t> deg(g.n);
  for (uint32_t v = 0; v < g.n; ++v) deg[v] = g.degree(v);
  if (!write_file(base + "_csr.deg", deg.data(), deg.size() * 4, err)) return false;
  if (write_wgh) {
    if (g.wgh.size() == E) {
      if (!write_file(base + "_csr.wgh", g.wgh.data(), E * 4, err)) return false;
    } else {
      std::vector<uint32_t> z(E, 0u);
      if (!write_file(base + "_csr.wgh", z.data(), E * 4, err)) return false;
    }
  }
  return true;
}

bool read_csr_files(const std::string& base, HostCsr* g, std::string* err) {
  const int64_t deg_bytes = file_size(base + "_csr.deg");
  const int64_t idx_bytes = file_size(base + "_csr.idx");
  const int64_t ngh_bytes = file_size(base + "_csr.ngh");
  if (deg_bytes < 0 || idx_bytes < 0 || ngh_bytes < 0) {
    if (err) *err = "missing " + base + "_csr.{idx,ngh,deg}";
    return false;
  }
  const uint64_t n = (uint64_t)deg_bytes / 4;
  const uint64_t E = (uint64_t)ngh_bytes / 4;
  if (n > UINT32_MAX) {
    if (err) *err = "more than 2^32 vertices";
    return false;
  }
  g->n = (uint32_t)n;
  g->offsets.assign(n + 1, 0);
  if ((uint64_t)idx_bytes == 8 * (n + 1)) {
    if (!read_file(base + "_csr.idx", g->offsets.data(), 8 * (n + 1), err)) return false;
  } else if ((uint64_t)idx_bytes == 4 * (n + 1)) {
    std::vector<uint32_t> o32(n + 1);
    if (!read_file(base + "_csr.idx", o32.data(), 4 * (n + 1), err)) return false;
    for (uint64_t i = 0; i <= n; ++i) g->offsets[i] = o32[i];
  } else {
    if (err) *err = base + "_csr.idx size matches neither u32 nor u64 offsets for V=" + std::to_string(n);
    return false;
  }
  if (g->offsets[n] != E) {
    if (err) *err = "row offsets end at " + std::to_string(g->offsets[n]) + " but " + std::to_string(E) + " neighbours are stored";
    return false;
  }
  for (uint64_t i = 0; i < n; ++i)
    if (g->offsets[i] > g->offsets[i + 1]) {
      if (err) *err = "row offsets decrease at vertex " + std::to_string(i);
      return false;
    }
  g->ngh.resize(E);
  if (!read_file(base + "_csr.ngh", g->ngh.data(), E * 4, err)) return false;
  for (uint64_t i = 0; i < E; ++i)
    if (g->ngh[i] >= n) {
      if (err) *err = "neighbour id out of range at edge " + std::to_string(i);
      return false;
    }
  const int64_t wgh_bytes = file_size(base + "_csr.wgh");
  g->wgh.clear();
  if (wgh_bytes == (int64_t)(E * 4)) {
    g->wgh.resize(E);
    if (!read_file(base + "_csr.wgh", g->wgh.data(), E * 4, err)) return false;
  }
  return true;
}

bool write_page_files(const std::string& base, const HostCsr& g, uint64_t P, bool weighted,
                      std::vector<scaph_page>* pages_out, std::string* err) {
  std::vector<scaph_page> pages;
  if (!page_cuts(g, P, weighted, &pages, err)) return false;
  std::unique_ptr<uint32_t[]> buf(new uint32_t[P]);
  for (size_t i = 0; i < pages.size(); ++i) {
    std::memset(buf.get(), 0, P * 4);
    fill_page(g, pages[i], weighted, buf.get());
    if (!write_file(base + "page" + std::to_string(i), buf.get(), P * 4, err)) return false;
  }
  if (!pages.empty() &&
      !write_file(base + "page.info", pages.data(), pages.size() * sizeof(scaph_page), err))
    return false;
  GraphInfo info;
  info.pagesize = (uint32_t)P;
  info.nodenum = g.n;
  info.edgenum = g.num_edges();
  if (!write_file(base + "graph.info", &info, sizeof(info), err)) return false;
  if (pages_out) *pages_out = pages;
  return true;
}

bool read_page_info(const std::string& base, std::vector<scaph_page>* pages, std::string* err) {
  const int64_t b = file_size(base + "page.info");
  if (b < 0 || b % sizeof(scaph_page)) {
    if (err) *err = "missing or malformed " + base + "page.info";
    return false;
  }
  pages->resize((size_t)b / sizeof(scaph_page));
  return read_file(base + "page.info", pages->data(), (size_t)b, err);
}

bool read_graph_info(const std::string& base, GraphInfo* info, std::string* err) {
  const int64_t b = file_size(base + "graph.info");
  if (b < 12) {
    if (err) *err = "missing or short " + base + "graph.info";
    return false;
  }
  std::memset(info, 0, sizeof(*info));
  // the engine reads the first 12 bytes (src/graph.h:79-87); read what exists
  return read_file(base + "graph.info", info, (size_t)std::min<int64_t>(b, 16), err);
}

bool read_page_file(const std::string& base, int index, uint64_t P, uint32_t* buf,
                    std::string* err) {
  const std::string path = base + "page" + std::to_string(index);
  const int64_t b = file_size(path);
  if (b != (int64_t)(P * 4)) {
    if (err) *err = path + " is not exactly pagesize words";
    return false;
  }
  return read_file(path, buf, P * 4, err);
}

HostCsr transpose(const HostCsr& g) {
  HostCsr t;
  t.n = g.n;
  t.offsets.assign((size_t)g.n + 1, 0);
  for (uint32_t w : g.ngh) t.offsets[(size_t)w + 1]++;
  for (uint32_t v = 0; v < g.n; ++v) t.offsets[v + 1] += t.offsets[v];
  t.ngh.resize(g.ngh.size());
  std::vector<uint64_t> cur(t.offsets.begin(), t.offsets.end() - 1);
  for (uint32_t u = 0; u < g.n; ++u)
    for (uint64_t k = g.offsets[u]; k < g.offsets[(size_t)u + 1]; ++k) t.ngh[cur[g.ngh[k]]++] = u;
  return t;
}

namespace {

// Buffered decimal text output (one large fwrite per 8 MiB).
class TextOut {
 public:
  explicit TextOut(const std::string& path) : f_(std::fopen(path.c_str(), "wb")) {
    buf_.reserve(kCap + 64);
  }
  ~TextOut() {
    if (f_) std::fclose(f_);
  }
  bool ok() const { return f_ != nullptr && good_; }
  void text(const char* s) {
    buf_.append(s);
    spill();
  }
  void num(uint64_t x, char end) {
    char tmp[24];
    const auto r = std::to_chars(tmp, tmp + sizeof tmp, x);
    buf_.append(tmp, r.ptr);
    buf_.push_back(end);
    spill();
  }
  bool close() {
    flush();
    const bool fine = f_ && good_ && std::fclose(f_) == 0;
    f_ = nullptr;
    return fine;
  }

 private:
  static constexpr size_t kCap = 8u << 20;
  void spill() {
    if (buf_.size() >= kCap) flush();
  }
  void flush() {
    if (f_ && !buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size())
      good_ = false;
    buf_.clear();
  }
  std::FILE* f_;
  std::string buf_;
  bool good_ = true;
};

bool export_check(const HostCsr& g, bool weighted, const std::string& path, TextOut& out,
                  std::string* err) {
  if (weighted && g.wgh.size() != g.ngh.size()) {
    if (err) *err = "weighted export needs the _csr.wgh file";
    return false;
  }
  if (!out.ok()) {
    if (err) *err = "cannot write " + path;
    return false;
  }
  return true;
}

}  // namespace

bool write_ligra(const std::string& base, const HostCsr& g, bool weighted, std::string* err) {
  const std::string path = base + (weighted ? ".ligra.weight" : ".ligra");
  TextOut out(path);
  if (!export_check(g, weighted, path, out, err)) return false;
  out.text(weighted ? "WeightedAdjacencyGraph\n" : "AdjacencyGraph\n");
  out.num(g.n, '\n');
  out.num(g.ngh.size(), '\n');
  for (uint32_t v = 0; v < g.n; ++v) out.num(g.offsets[v], '\n');
  for (uint32_t x : g.ngh) out.num(x, '\n');
  if (weighted)
    for (uint32_t w : g.wgh) out.num(w, '\n');
  if (!out.close()) {
    if (err) *err = "write failed: " + path;
    return false;
  }
  return true;
}

bool write_totem(const std::string& base, const HostCsr& g, bool weighted, std::string* err) {
  const std::string path = base + (weighted ? "weight.totem" : ".totem");
  TextOut out(path);
  if (!export_check(g, weighted, path, out, err)) return false;
  out.text("#Nodes: ");
  out.num(g.n, '\n');
  out.text("#Edges: ");
  out.num(g.ngh.size(), '\n');
  out.text("#Undirected\n");
  for (uint32_t u = 0; u < g.n; ++u)
    for (uint64_t k = g.offsets[u]; k < g.offsets[(size_t)u + 1]; ++k) {
      out.num(u, '\t');
      if (weighted) {
        out.num(g.ngh[k], '\t');
        out.num(g.wgh[k], '\n');
      } else {
        out.num(g.ngh[k], '\n');
      }
    }
  if (!out.close()) {
    if (err) *err = "write failed: " + path;
    return false;
  }
  return true;
}

uint64_t edge_multiset_digest(const HostCsr& g, bool transposed) {
  auto mix = [](uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  uint64_t sum = 0;
  for (uint32_t u = 0; u < g.n; ++u)
    for (uint64_t k = g.offsets[u]; k < g.offsets[(size_t)u + 1]; ++k) {
      const uint64_t a = transposed ? g.ngh[k] : u, b = transposed ? u : g.ngh[k];
      sum += mix((a << 32 | b) + 0x9E3779B97F4A7C15ull);
    }
  return sum;
}

bool write_transposed_files(const std::string& base, const HostCsr& g, uint64_t pagesize_words,
                            std::vector<scaph_page>* pages_out, std::string* err) {
  const HostCsr t = transpose(g);
  return write_csr_files(base + "_in", t, true, false, err) &&
         write_page_files(base + "_in", t, pagesize_words, false, pages_out, err);
}

std::vector<uint32_t> select_hubs(const std::vector<uint32_t>& deg, uint32_t max_hubs) {
  std::vector<uint32_t> ids;
  for (uint32_t v = 0; v < deg.size(); ++v)
    if (deg[v] > 0) ids.push_back(v);
  auto better = [&](uint32_t a, uint32_t b) { return deg[a] != deg[b] ? deg[a] > deg[b] : a < b; };
  const size_t k = std::min<size_t>(ids.size(), max_hubs);
  std::partial_sort(ids.begin(), ids.begin() + k, ids.end(), better);
  ids.resize(k);
  return ids;
}

void encode_hub_in_lists(HostCsr* in, const std::vector<uint32_t>& hubs) {
  std::vector<uint32_t> idx(in->n, 0xFFFFFFFFu);
  for (uint32_t i = 0; i < hubs.size(); ++i) idx[hubs[i]] = i;
  std::vector<uint32_t> hub_part, rest;
  for (uint32_t v = 0; v < in->n; ++v) {
    hub_part.clear();
    rest.clear();
    for (uint64_t k = in->offsets[v]; k < in->offsets[(size_t)v + 1]; ++k) {
      const uint32_t u = in->ngh[k];
      if (idx[u] != 0xFFFFFFFFu) hub_part.push_back(idx[u]);
      else rest.push_back(u);
    }
    std::sort(hub_part.begin(), hub_part.end());
    std::sort(rest.begin(), rest.end());
    uint64_t k = in->offsets[v];
    for (uint32_t h : hub_part) in->ngh[k++] = 0x80000000u | h;
    for (uint32_t u : rest) in->ngh[k++] = u;
  }
}

}  // namespace scaph

// Bookshelf reader (host code).
//
// Reads the UCLA Bookshelf files named by an .aux file (.nodes, .nets, .wts,
// .scl, .pl; .shapes and .route are not read) into the arrays the placer
// consumes, with the node, net and pin orders of the reference's PlaceDB:
//   * pins of a net sorted by (node name, pin name); one pin kept per node;
//     an output pin is swapped to the front of its net as it is added
//     (dreamplace/ops/place_io/src/PlaceDB.cpp:973-1068, 1485-1510);
//   * pin offsets from the node's lower-left corner:
//     round(offset + size / 2) (PlaceDB.cpp:1053-1061);
//   * positions round(x), round(y); status from the .pl flag (/FIXED,
//     /FIXED_NI = IO pin, none = placed); a movable node taller than 5 rows
//     becomes dummy-fixed (PlaceDB.cpp:1109-1164, utility/src/defs.h:74);
//   * nodes ordered unplaced, placed, dummy-fixed, fixed, IO pins, ties by
//     file order (PlaceDB.cpp:2017-2090);
//   * nets ordered by degree, ties by file order, then pins by net, ties by
//     creation order (PlaceDB.cpp:1854-1990);
//   * rows sorted bottom-up; die = bounding box of the rows
//     (PlaceDB.cpp:819-820, PyPlaceDB.cpp:434-440);
//   * fixed nodes and IO pins in another orientation than N: size and pin
//     offsets restated for N (PyPlaceDB.cpp:669-740).
// The Bookshelf grammar itself follows the published format (the reference
// parses it with the Limbo library, a submodule that is not vendored).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <numeric>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dreamplace_amd.h"

namespace {

enum Status { UNPLACED = 0, PLACED = 1, DUMMY_FIXED = 2, FIXED = 3 };

struct BNode {
  std::string name;
  int64_t w = 0, h = 0, x = 0, y = 0;
  int orient = DP_ORIENT_N;
  int status = UNPLACED;
  bool io = false;  // terminal_NI / FIXED_NI
  std::vector<int64_t> pins;
};

struct BPin {
  int64_t node = 0, net = 0;
  double ox = 0, oy = 0;  // offset from the lower-left corner
  int direct = DP_PIN_INOUT;
};

struct BNet {
  std::string name;
  double weight = 1.0;
  bool ignore = false;
  std::vector<int64_t> pins;
};

struct BRow {
  double xl, yl, xh, yh, site_width, height;
};

struct Design {
  std::vector<BNode> nodes;
  std::vector<BPin> pins;
  std::vector<BNet> nets;
  std::vector<BRow> rows;
  std::unordered_map<std::string, int64_t> node_index, net_index;
  double row_height = 0, site_width = 0;
  int64_t num_terminal_nis = 0;
  std::string error;
};

// -- tokenising ---------------------------------------------------------------
bool read_file(const std::string& path, std::string* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  *out = ss.str();
  return true;
}

// Streams the non-empty lines of a file, comments (#...) removed, as
// whitespace-separated tokens (':' is a token of its own).
struct LineReader {
  const std::string& text;
  size_t pos = 0;
  std::vector<std::string> toks;
  explicit LineReader(const std::string& t) : text(t) {}
  bool next() {
    const size_t n = text.size();
    while (pos < n) {
      size_t e = text.find('\n', pos);
      if (e == std::string::npos) e = n;
      toks.clear();
      size_t j = pos;
      while (j < e) {
        char c = text[j];
        if (c == '#') break;
        if (c == ' ' || c == '\t' || c == '\r') {
          ++j;
          continue;
        }
        if (c == ':') {
          toks.emplace_back(":");
          ++j;
          continue;
        }
        size_t k = j;
        while (k < e && text[k] != ' ' && text[k] != '\t' && text[k] != '\r' && text[k] != ':' &&
               text[k] != '#')
          ++k;
        toks.emplace_back(text, j, k - j);
        j = k;
      }
      pos = e + 1;
      if (!toks.empty()) return true;
    }
    return false;
  }
};

bool is_header(const std::vector<std::string>& t) { return t.size() >= 2 && t[0] == "UCLA"; }

bool iequals(const std::string& a, const char* b) {
  size_t n = strlen(b);
  if (a.size() != n) return false;
  for (size_t i = 0; i < n; ++i)
    if (tolower((unsigned char)a[i]) != tolower((unsigned char)b[i])) return false;
  return true;
}

int orient_code(const std::string& s) {
  static const char* names[] = {"N", "S", "W", "E", "FN", "FS", "FW", "FE"};
  for (int i = 0; i < 8; ++i)
    if (iequals(s, names[i])) return i;
  return DP_ORIENT_UNKNOWN;
}

bool to_double(const std::string& s, double* v) {
  char* end = nullptr;
  *v = strtod(s.c_str(), &end);
  return end && *end == 0 && !s.empty();
}

// -- files ----------------------------------------------------------------------
bool parse_nodes(Design& d, const std::string& text) {
  LineReader L(text);
  while (L.next()) {
    auto& t = L.toks;
    if (is_header(t)) continue;
    if (t.size() >= 3 && t[1] == ":" && (t[0] == "NumNodes" || t[0] == "NumTerminals")) continue;
    if (t.size() < 3) {
      d.error = ".nodes: bad line starting '" + t[0] + "'";
      return false;
    }
    double w, h;
    if (!to_double(t[1], &w) || !to_double(t[2], &h)) {
      d.error = ".nodes: bad size for " + t[0];
      return false;
    }
    if (d.node_index.count(t[0])) continue;  // duplicate: first one kept
    BNode nd;
    nd.name = t[0];
    nd.w = (int64_t)w;
    nd.h = (int64_t)h;
    if (t.size() >= 4 && iequals(t[3], "terminal_NI")) {
      nd.io = true;
      d.num_terminal_nis += 1;
    }
    d.node_index.emplace(nd.name, (int64_t)d.nodes.size());
    d.nodes.push_back(std::move(nd));
  }
  return true;
}

struct NetPinRec {
  std::string node, pin;
  char direct;
  double ox, oy;
};

void add_net(Design& d, const std::string& name, std::vector<NetPinRec>& np) {
  std::sort(np.begin(), np.end(), [](const NetPinRec& a, const NetPinRec& b) {
    return a.node < b.node || (a.node == b.node && a.pin < b.pin);
  });
  np.erase(std::unique(np.begin(), np.end(),
                       [](const NetPinRec& a, const NetPinRec& b) { return a.node == b.node; }),
           np.end());
  bool ignore = np.size() < 2;
  bool one_node = true;
  for (size_t i = 1; i < np.size(); ++i)
    if (np[i - 1].node != np[i].node) one_node = false;
  if (one_node) ignore = true;
  if (d.net_index.count(name)) return;  // duplicate net: first one kept
  int64_t id = (int64_t)d.nets.size();
  d.net_index.emplace(name, id);
  d.nets.emplace_back();
  BNet& net = d.nets.back();
  net.name = name;
  net.ignore = ignore;
  for (auto& p : np) {
    auto f = d.node_index.find(p.node);
    if (f == d.node_index.end()) continue;  // unknown node: pin skipped
    BNode& node = d.nodes[f->second];
    BPin pin;
    pin.node = f->second;
    pin.net = id;
    pin.direct = p.direct == 'I' ? DP_PIN_INPUT : p.direct == 'O' ? DP_PIN_OUTPUT : DP_PIN_INOUT;
    pin.ox = std::round(p.ox + node.w * 0.5);
    pin.oy = std::round(p.oy + node.h * 0.5);
    int64_t pid = (int64_t)d.pins.size();
    d.pins.push_back(pin);
    node.pins.push_back(pid);
    net.pins.push_back(pid);
    if (pin.direct == DP_PIN_OUTPUT) std::swap(net.pins.front(), net.pins.back());
  }
}

bool parse_nets(Design& d, const std::string& text) {
  LineReader L(text);
  std::vector<NetPinRec> cur;
  std::string name;
  int64_t want = -1, count = 0;
  auto flush = [&]() {
    if (want >= 0) add_net(d, name, cur);
    cur.clear();
    want = -1;
  };
  while (L.next()) {
    auto& t = L.toks;
    if (is_header(t)) continue;
    if (t[0] == "NumNets" || t[0] == "NumPins") continue;
    if (t[0] == "NetDegree") {
      flush();
      if (t.size() < 3 || t[1] != ":") {
        d.error = ".nets: bad NetDegree line";
        return false;
      }
      want = strtoll(t[2].c_str(), nullptr, 10);
      name = t.size() >= 4 ? t[3] : ("net_" + std::to_string(count));
      ++count;
      continue;
    }
    if (want < 0) {
      d.error = ".nets: pin line outside a net";
      return false;
    }
    // node [direction] [: ox oy]
    NetPinRec r{t[0], "", 'B', 0.0, 0.0};
    size_t k = 1;
    if (k < t.size() && t[k] != ":") {
      r.direct = t[k][0];
      ++k;
    }
    if (k < t.size() && t[k] == ":") {
      if (k + 2 >= t.size() || !to_double(t[k + 1], &r.ox) || !to_double(t[k + 2], &r.oy)) {
        d.error = ".nets: bad pin offset for " + t[0];
        return false;
      }
      if (k + 3 < t.size()) r.pin = t[k + 3];
    }
    cur.push_back(r);
  }
  flush();
  return true;
}

bool parse_wts(Design& d, const std::string& text) {
  LineReader L(text);
  while (L.next()) {
    auto& t = L.toks;
    if (is_header(t) || t.size() < 2) continue;
    auto f = d.net_index.find(t[0]);
    double w;
    if (f != d.net_index.end() && to_double(t[1], &w)) d.nets[f->second].weight = w;
  }
  return true;
}

bool parse_scl(Design& d, const std::string& text) {
  LineReader L(text);
  BRow row{0, 0, 0, 0, 1, 0};
  double origin = 0, sites = 0, coord = 0, height = 0, sw = 1;
  bool in_row = false;
  while (L.next()) {
    auto& t = L.toks;
    if (is_header(t)) continue;
    if (t[0] == "NumRows") continue;
    if (t[0] == "CoreRow") {
      in_row = true;
      origin = sites = coord = height = 0;
      sw = 1;
      continue;
    }
    if (!in_row) continue;
    if (t[0] == "End") {
      row.xl = origin;
      row.yl = coord;
      row.xh = origin + sw * sites;
      row.yh = coord + height;
      row.site_width = sw;
      row.height = height;
      d.rows.push_back(row);
      d.row_height = height;
      d.site_width = sw;
      in_row = false;
      continue;
    }
    // "key : value" pairs, possibly two on one line
    for (size_t k = 0; k + 2 < t.size(); k += 3) {
      if (t[k + 1] != ":") break;
      double v;
      if (!to_double(t[k + 2], &v)) break;
      const std::string& key = t[k];
      if (iequals(key, "Coordinate")) coord = v;
      else if (iequals(key, "Height")) height = v;
      else if (iequals(key, "Sitewidth")) sw = v;
      else if (iequals(key, "SubrowOrigin")) origin = v;
      else if (iequals(key, "NumSites")) sites = v;
    }
  }
  std::stable_sort(d.rows.begin(), d.rows.end(),
                   [](const BRow& a, const BRow& b) { return a.yl < b.yl; });
  return true;
}

bool parse_pl(Design& d, const std::string& text) {
  LineReader L(text);
  while (L.next()) {
    auto& t = L.toks;
    if (is_header(t)) continue;
    if (t.size() < 3) continue;
    auto f = d.node_index.find(t[0]);
    if (f == d.node_index.end()) continue;
    double x, y;
    if (!to_double(t[1], &x) || !to_double(t[2], &y)) {
      d.error = ".pl: bad position for " + t[0];
      return false;
    }
    BNode& nd = d.nodes[f->second];
    nd.x = (int64_t)std::round(x);
    nd.y = (int64_t)std::round(y);
    size_t k = 3;
    if (k < t.size() && t[k] == ":") ++k;
    if (k < t.size() && t[k][0] != '/') nd.orient = orient_code(t[k++]);
    std::string status = k < t.size() ? t[k] : "";
    if (!status.empty() && status[0] == '/') status = status.substr(1);
    if (status.empty()) {
      nd.status = PLACED;
    } else if (iequals(status, "FIXED_NI") || iequals(status, "FIXED")) {
      // an IO pin is known by its .nodes type (terminal_NI); FIXED_NI only fixes it
      nd.status = FIXED;
    } else if (iequals(status, "PLACED")) {
      nd.status = PLACED;
    } else if (iequals(status, "UNPLACED")) {
      nd.status = UNPLACED;
    } else {
      nd.status = PLACED;
    }
    if (nd.status != FIXED && d.row_height > 0 && nd.h > d.row_height * 5) nd.status = DUMMY_FIXED;
  }
  return true;
}

// -- orders -----------------------------------------------------------------
// Orientation as (clockwise degrees, mirrored about the y axis); UNKNOWN
// counts as N (PyPlaceDB.cpp:545-594).
void orient_degree_flip(int orient, int* degree, int* flip) {
  static const int deg[8] = {0, 180, 90, 270, 0, 180, 90, 270};
  *degree = orient >= 0 && orient < 8 ? deg[orient] : 0;
  *flip = orient >= DP_ORIENT_FN && orient <= DP_ORIENT_FE ? 1 : 0;
}

// Fixed nodes and IO pins placed in an orientation other than N get their
// size and their pins' offsets restated for N (PyPlaceDB::convertOrient,
// PyPlaceDB.cpp:669-740, with the rotations of :596-667): rotate by
// (360 - degree) % 360, then mirror about the y axis if the orientation is
// flipped.  The recorded orientation is kept.  Movable nodes are placed in
// N anyway and are left alone.
void convert_orient(Design& d) {
  for (auto& nd : d.nodes) {
    if (nd.status != FIXED || nd.orient == DP_ORIENT_N || nd.orient == DP_ORIENT_UNKNOWN) continue;
    int degree, flip;
    orient_degree_flip(nd.orient, &degree, &flip);
    const int rot = (360 - degree) % 360;
    const double w = (double)nd.w, h = (double)nd.h;
    for (int64_t p : nd.pins) {
      BPin& pin = d.pins[p];
      const double x = pin.ox, y = pin.oy;
      if (rot == 180) {
        pin.ox = w - x;
        pin.oy = h - y;
      } else if (rot == 270) {
        pin.ox = h - y;
        pin.oy = x;
      } else if (rot == 90) {
        pin.ox = y;
        pin.oy = w - x;
      }
    }
    if (rot == 90 || rot == 270) std::swap(nd.w, nd.h);
    if (flip)
      for (int64_t p : nd.pins) d.pins[p].ox = (double)nd.w - d.pins[p].ox;
  }
}

void sort_nodes(Design& d) {
  const int64_t n = (int64_t)d.nodes.size();
  auto key = [&](int64_t i) {
    const BNode& nd = d.nodes[i];
    int64_t k = nd.status == FIXED ? 512 : nd.status == DUMMY_FIXED ? 64 : nd.status == PLACED ? 8 : 1;
    if (nd.io) k += 2048;
    return k;
  };
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key(a) < key(b); });
  std::vector<int64_t> where(n);
  for (int64_t i = 0; i < n; ++i) where[order[i]] = i;
  std::vector<BNode> nodes(n);
  for (int64_t i = 0; i < n; ++i) nodes[i] = std::move(d.nodes[order[i]]);
  d.nodes.swap(nodes);
  for (auto& p : d.pins) p.node = where[p.node];
  for (auto& kv : d.node_index) kv.second = where[kv.second];
}

void sort_nets(Design& d) {
  const int64_t E = (int64_t)d.nets.size(), P = (int64_t)d.pins.size();
  std::vector<int64_t> order(E);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return d.nets[a].pins.size() < d.nets[b].pins.size();
  });
  std::vector<int64_t> where(E);
  for (int64_t i = 0; i < E; ++i) where[order[i]] = i;
  std::vector<BNet> nets(E);
  for (int64_t i = 0; i < E; ++i) nets[i] = std::move(d.nets[order[i]]);
  d.nets.swap(nets);
  for (auto& kv : d.net_index) kv.second = where[kv.second];
  for (auto& p : d.pins) p.net = where[p.net];
  // pins by net, ties by creation order
  std::vector<int64_t> porder(P);
  std::iota(porder.begin(), porder.end(), 0);
  std::stable_sort(porder.begin(), porder.end(),
                   [&](int64_t a, int64_t b) { return d.pins[a].net < d.pins[b].net; });
  std::vector<int64_t> pwhere(P);
  for (int64_t i = 0; i < P; ++i) pwhere[porder[i]] = i;
  std::vector<BPin> pins(P);
  for (int64_t i = 0; i < P; ++i) pins[i] = d.pins[porder[i]];
  d.pins.swap(pins);
  for (auto& nd : d.nodes)
    for (auto& p : nd.pins) p = pwhere[p];
  for (auto& net : d.nets)
    for (auto& p : net.pins) p = pwhere[p];
}

std::string dir_of(const std::string& path) {
  size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

}  // namespace

namespace dp {
void set_error(const char* fmt, ...);  // the library's per-thread last error
}

extern "C" {

int dp_bookshelf_read(const char* aux_path, int sort_nets_by_degree, void** handle) {
  if (!aux_path || !handle) {
    dp::set_error("dp_bookshelf_read: null argument");
    return DP_ERR_INVALID_ARGUMENT;
  }
  *handle = nullptr;
  std::string aux;
  if (!read_file(aux_path, &aux)) {
    dp::set_error("cannot read %s", aux_path);
    return DP_ERR_INVALID_ARGUMENT;
  }
  std::string base = dir_of(aux_path);
  std::string f_nodes, f_nets, f_wts, f_pl, f_scl;
  LineReader L(aux);
  while (L.next())
    for (auto& tok : L.toks) {
      auto ends = [&](const char* ext) {
        size_t n = strlen(ext);
        return tok.size() > n && tok.compare(tok.size() - n, n, ext) == 0;
      };
      std::string p = tok[0] == '/' ? tok : base + "/" + tok;
      if (ends(".nodes")) f_nodes = p;
      else if (ends(".nets")) f_nets = p;
      else if (ends(".wts")) f_wts = p;
      else if (ends(".pl")) f_pl = p;
      else if (ends(".scl")) f_scl = p;
    }
  if (f_nodes.empty() || f_nets.empty() || f_pl.empty() || f_scl.empty()) {
    dp::set_error("%s must name .nodes, .nets, .pl and .scl files", aux_path);
    return DP_ERR_INVALID_ARGUMENT;
  }
  Design* d = new Design();
  std::string text;
  bool ok = true;
  auto step = [&](const std::string& path, bool (*fn)(Design&, const std::string&), bool need) {
    if (!ok) return;
    if (path.empty()) return;
    if (!read_file(path, &text)) {
      if (need) {
        d->error = "cannot read " + path;
        ok = false;
      }
      return;
    }
    ok = fn(*d, text);
  };
  // rows before positions: the dummy-fixed test needs the row height
  step(f_nodes, parse_nodes, true);
  step(f_nets, parse_nets, true);
  step(f_wts, parse_wts, false);
  step(f_scl, parse_scl, true);
  step(f_pl, parse_pl, true);
  if (ok && d->rows.empty()) {
    d->error = ".scl holds no rows";
    ok = false;
  }
  if (!ok) {
    dp::set_error("%s", d->error.c_str());
    delete d;
    return DP_ERR_INVALID_ARGUMENT;
  }
  sort_nodes(*d);
  if (sort_nets_by_degree) sort_nets(*d);
  convert_orient(*d);
  *handle = d;
  return DP_OK;
}

int dp_bookshelf_sizes(void* handle, int64_t* sizes) {
  if (!handle || !sizes) return DP_ERR_INVALID_ARGUMENT;
  const Design& d = *(const Design*)handle;
  int64_t fixed = 0, movable_pins = 0, node_bytes = 0, net_bytes = 0;
  for (auto& nd : d.nodes) {
    if (nd.status == FIXED && !nd.io) ++fixed;
    node_bytes += (int64_t)nd.name.size() + 1;
  }
  for (auto& p : d.pins)
    if (d.nodes[p.node].status != FIXED) ++movable_pins;
  for (auto& net : d.nets) net_bytes += (int64_t)net.name.size() + 1;
  int64_t io = 0;
  for (auto& nd : d.nodes) io += nd.io ? 1 : 0;
  sizes[0] = (int64_t)d.nodes.size();
  sizes[1] = fixed;                     // num_terminals
  sizes[2] = io;                        // num_terminal_NIs
  sizes[3] = (int64_t)d.nets.size();
  sizes[4] = (int64_t)d.pins.size();
  sizes[5] = (int64_t)d.rows.size();
  sizes[6] = node_bytes;
  sizes[7] = net_bytes;
  sizes[8] = movable_pins;
  return DP_OK;
}

int dp_bookshelf_export(void* handle, double* node_x, double* node_y, double* node_size_x,
                        double* node_size_y, int32_t* node_orient, int32_t* node_status,
                        char* node_names, double* pin_offset_x, double* pin_offset_y,
                        int32_t* pin_direct, int32_t* pin2node_map, int32_t* pin2net_map,
                        int32_t* flat_net2pin_map, int32_t* flat_net2pin_start_map,
                        int32_t* flat_node2pin_map, int32_t* flat_node2pin_start_map,
                        double* net_weights, uint8_t* net_ignore, char* net_names, double* rows,
                        double* die) {
  if (!handle) return DP_ERR_INVALID_ARGUMENT;
  const Design& d = *(const Design*)handle;
  int64_t c = 0, o = 0;
  for (size_t i = 0; i < d.nodes.size(); ++i) {
    const BNode& nd = d.nodes[i];
    if (node_x) node_x[i] = (double)nd.x;
    if (node_y) node_y[i] = (double)nd.y;
    if (node_size_x) node_size_x[i] = (double)nd.w;
    if (node_size_y) node_size_y[i] = (double)nd.h;
    if (node_orient) node_orient[i] = nd.orient;
    if (node_status) node_status[i] = nd.status;
    if (node_names) {
      memcpy(node_names + o, nd.name.c_str(), nd.name.size() + 1);
      o += (int64_t)nd.name.size() + 1;
    }
    if (flat_node2pin_start_map) flat_node2pin_start_map[i] = (int32_t)c;
    for (int64_t p : nd.pins) {
      if (flat_node2pin_map) flat_node2pin_map[c] = (int32_t)p;
      ++c;
    }
  }
  if (flat_node2pin_start_map) flat_node2pin_start_map[d.nodes.size()] = (int32_t)c;
  for (size_t i = 0; i < d.pins.size(); ++i) {
    const BPin& p = d.pins[i];
    if (pin_offset_x) pin_offset_x[i] = p.ox;
    if (pin_offset_y) pin_offset_y[i] = p.oy;
    if (pin_direct) pin_direct[i] = p.direct;
    if (pin2node_map) pin2node_map[i] = (int32_t)p.node;
    if (pin2net_map) pin2net_map[i] = (int32_t)p.net;
  }
  c = 0;
  o = 0;
  for (size_t i = 0; i < d.nets.size(); ++i) {
    const BNet& net = d.nets[i];
    if (flat_net2pin_start_map) flat_net2pin_start_map[i] = (int32_t)c;
    for (int64_t p : net.pins) {
      if (flat_net2pin_map) flat_net2pin_map[c] = (int32_t)p;
      ++c;
    }
    if (net_weights) net_weights[i] = net.weight;
    if (net_ignore) net_ignore[i] = net.ignore ? 1 : 0;
    if (net_names) {
      memcpy(net_names + o, net.name.c_str(), net.name.size() + 1);
      o += (int64_t)net.name.size() + 1;
    }
  }
  if (flat_net2pin_start_map) flat_net2pin_start_map[d.nets.size()] = (int32_t)c;
  double xl = 0, yl = 0, xh = 0, yh = 0;
  for (size_t i = 0; i < d.rows.size(); ++i) {
    const BRow& r = d.rows[i];
    if (rows) {
      rows[4 * i + 0] = r.xl;
      rows[4 * i + 1] = r.yl;
      rows[4 * i + 2] = r.xh;
      rows[4 * i + 3] = r.yh;
    }
    if (i == 0) {
      xl = r.xl, yl = r.yl, xh = r.xh, yh = r.yh;
    } else {
      xl = std::min(xl, r.xl), yl = std::min(yl, r.yl);
      xh = std::max(xh, r.xh), yh = std::max(yh, r.yh);
    }
  }
  if (die) {
    die[0] = xl, die[1] = yl, die[2] = xh, die[3] = yh;
    die[4] = d.row_height;
    die[5] = d.site_width;
  }
  return DP_OK;
}

void dp_bookshelf_free(void* handle) { delete (Design*)handle; }

// Area of the union of n rectangles clipped to `clip` (xl, yl, xh, yh;
// nullable = no clip): a sweep over x with a segment tree of covered y
// length (the polygon-set area of PyPlaceDB::computeAreaStatistics).
double dp_rect_union_area(int64_t n, const double* xl, const double* yl, const double* xh,
                          const double* yh, const double* clip) {
  struct Ev {
    double x;
    int64_t lo, hi;
    int d;
  };
  std::vector<double> ys;
  std::vector<double> r(4 * (size_t)std::max<int64_t>(n, 0));
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    double a = xl[i], b = yl[i], c = xh[i], e = yh[i];
    if (clip) {
      a = std::max(a, clip[0]), b = std::max(b, clip[1]);
      c = std::min(c, clip[2]), e = std::min(e, clip[3]);
    }
    if (!(c > a && e > b)) continue;
    r[4 * m] = a, r[4 * m + 1] = b, r[4 * m + 2] = c, r[4 * m + 3] = e;
    ys.push_back(b);
    ys.push_back(e);
    ++m;
  }
  if (m == 0) return 0.0;
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  const int64_t K = (int64_t)ys.size() - 1;  // elementary intervals
  std::vector<Ev> ev;
  ev.reserve(2 * m);
  for (int64_t i = 0; i < m; ++i) {
    int64_t lo = std::lower_bound(ys.begin(), ys.end(), r[4 * i + 1]) - ys.begin();
    int64_t hi = std::lower_bound(ys.begin(), ys.end(), r[4 * i + 3]) - ys.begin();
    ev.push_back({r[4 * i], lo, hi, +1});
    ev.push_back({r[4 * i + 2], lo, hi, -1});
  }
  std::sort(ev.begin(), ev.end(), [](const Ev& a, const Ev& b) { return a.x < b.x; });
  std::vector<int> cnt(4 * K, 0);
  std::vector<double> len(4 * K, 0.0);
  // add d over elementary intervals [lo, hi) within node covering [l, r)
  std::function<void(int64_t, int64_t, int64_t, int64_t, int64_t, int)> upd =
      [&](int64_t node, int64_t l, int64_t rr, int64_t lo, int64_t hi, int d) {
        if (hi <= l || rr <= lo) return;
        if (lo <= l && rr <= hi) {
          cnt[node] += d;
        } else {
          int64_t mid = (l + rr) / 2;
          upd(2 * node + 1, l, mid, lo, hi, d);
          upd(2 * node + 2, mid, rr, lo, hi, d);
        }
        if (cnt[node] > 0) len[node] = ys[rr] - ys[l];
        else if (rr - l == 1) len[node] = 0;
        else len[node] = len[2 * node + 1] + len[2 * node + 2];
      };
  double area = 0, px = ev[0].x;
  for (const Ev& e : ev) {
    area += len[0] * (e.x - px);
    px = e.x;
    upd(0, 0, K, e.lo, e.hi, e.d);
  }
  return area;
}

}  // extern "C"

// Bookshelf (.aux/.nodes/.nets/.pl/.scl) reader and .pl writer for the C++
// host, following the reference parser's semantics (Parser/parser.cpp:44-820):
// nodes vs terminal / terminal_NI, pins sorted by net in file order with
// offsets from the module centre, rows in file order, and .pl lower-left
// corners clamped into the core for movable nodes (placedb.cpp:200-240).
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/eplace_amd.h"

namespace eplace_amd {

struct BookshelfDesign {
  std::string name, dir;
  std::vector<std::string> node_names, term_names;
  std::vector<float> node_w, node_h, node_x, node_y;
  std::vector<float> term_w, term_h, term_x, term_y;
  std::vector<uint8_t> term_ni;
  std::vector<int64_t> net_start{0};
  std::vector<int32_t> pin_obj;
  std::vector<float> pin_dx, pin_dy;
  std::vector<double> row_bottom, row_height, row_x0, row_x1;
  double site_width = 1.0;

  ep_design view() const {
    ep_design d{};
    d.num_nodes = (int32_t)node_w.size();
    d.num_terminals = (int32_t)term_w.size();
    d.num_nets = (int32_t)net_start.size() - 1;
    d.num_rows = (int32_t)row_bottom.size();
    d.num_pins = (int64_t)pin_obj.size();
    d.node_w = node_w.data();
    d.node_h = node_h.data();
    d.node_x = node_x.data();
    d.node_y = node_y.data();
    d.term_w = term_w.data();
    d.term_h = term_h.data();
    d.term_x = term_x.data();
    d.term_y = term_y.data();
    d.term_ni = term_ni.data();
    d.net_start = net_start.data();
    d.pin_obj = pin_obj.data();
    d.pin_dx = pin_dx.data();
    d.pin_dy = pin_dy.data();
    d.row_bottom = row_bottom.data();
    d.row_height = row_height.data();
    d.row_x0 = row_x0.data();
    d.row_x1 = row_x1.data();
    d.common_row_height = row_height.empty() ? 1.0 : row_height.back();
    return d;
  }

  void core(float &llx, float &lly, float &urx, float &ury) const {
    lly = (float)row_bottom.front();
    ury = (float)(row_bottom.back() + row_height.back());
    llx = (float)row_x0.front();
    urx = (float)row_x1.front();
    for (size_t r = 0; r < row_bottom.size(); ++r) {
      llx = std::min(llx, (float)row_x0[r]);
      urx = std::max(urx, (float)row_x1[r]);
    }
  }
};

namespace detail {
inline std::vector<std::string> tokens(const std::string &line) {
  std::vector<std::string> t;
  std::istringstream is(line);
  std::string s;
  while (is >> s) t.push_back(s);
  return t;
}
inline std::ifstream open(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  return f;
}
}  // namespace detail

inline void load_pl(BookshelfDesign &d, const std::string &path) {
  std::unordered_map<std::string, int> nidx, tidx;
  for (size_t i = 0; i < d.node_names.size(); ++i) nidx[d.node_names[i]] = (int)i;
  for (size_t i = 0; i < d.term_names.size(); ++i) tidx[d.term_names[i]] = (int)i;
  float llx, lly, urx, ury;
  d.core(llx, lly, urx, ury);
  std::ifstream f = detail::open(path);
  std::string line;
  std::getline(f, line);
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    for (char &ch : line)
      if (ch == '(' || ch == ')' || ch == ',' || ch == ':' || ch == '=' || ch == '\r') ch = ' ';
    auto t = detail::tokens(line);
    if (t.size() < 3) continue;
    float x = (float)atof(t[1].c_str()), y = (float)atof(t[2].c_str());
    auto it = nidx.find(t[0]);
    if (it != nidx.end()) {
      int i = it->second;
      float w = d.node_w[i], h = d.node_h[i];
      if (x < llx) x = llx;
      if (x + w > urx) x = (float)((double)(urx - w) - 1.0e-15);
      if (y < lly) y = lly;
      if (y + h > ury) y = (float)((double)(ury - h) - 1.0e-15);
      d.node_x[i] = x + (float)0.5 * w;
      d.node_y[i] = y + (float)0.5 * h;
      continue;
    }
    auto jt = tidx.find(t[0]);
    if (jt == tidx.end()) throw std::runtime_error("module " + t[0] + " in .pl is not in the design");
    int i = jt->second;
    d.term_x[i] = x + (float)0.5 * d.term_w[i];
    d.term_y[i] = y + (float)0.5 * d.term_h[i];
  }
}

inline BookshelfDesign read_bookshelf(const std::string &aux) {
  BookshelfDesign d;
  size_t slash = aux.rfind('/');
  d.dir = slash == std::string::npos ? std::string("./") : aux.substr(0, slash + 1);
  std::string base = aux.substr(slash == std::string::npos ? 0 : slash + 1);
  d.name = base.size() > 4 && base.substr(base.size() - 4) == ".aux" ? base.substr(0, base.size() - 4) : base;
  std::string nodes, nets, pl, scl;
  {
    std::ifstream f = detail::open(aux);
    std::string all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    for (char &ch : all)
      if (ch == ':') ch = ' ';
    for (const std::string &t : detail::tokens(all)) {
      auto ends = [&](const char *e) { return t.size() > strlen(e) && t.compare(t.size() - strlen(e), strlen(e), e) == 0; };
      if (ends(".nodes")) nodes = d.dir + t;
      else if (ends(".nets")) nets = d.dir + t;
      else if (ends(".pl")) pl = d.dir + t;
      else if (ends(".scl")) scl = d.dir + t;
    }
  }
  {  // .scl
    std::ifstream f = detail::open(scl);
    std::string line;
    std::getline(f, line);
    double site = 1.0;
    while (std::getline(f, line)) {
      auto t = detail::tokens(line);
      if (t.empty()) continue;
      if (t[0] == "CoreRow") {
        d.row_bottom.push_back(0);
        d.row_height.push_back(0);
        d.row_x0.push_back(0);
        d.row_x1.push_back(0);
      } else if (t[0] == "Coordinate" && t.size() >= 3) {
        d.row_bottom.back() = atof(t[2].c_str());
      } else if (t[0] == "Height" && t.size() >= 3) {
        d.row_height.back() = atof(t[2].c_str());
      } else if (t[0] == "Sitewidth" && t.size() >= 3) {
        site = atof(t[2].c_str());
        d.site_width = site;
      } else if (t.size() >= 6 && (t[3] == "Numsites" || t[3] == "NumSites")) {
        double origin = atof(t[2].c_str()), n = atof(t[5].c_str());
        d.row_x0.back() = origin;
        d.row_x1.back() = n * site + origin;
      }
    }
    if (d.row_bottom.empty()) throw std::runtime_error("no rows in " + scl);
  }
  std::unordered_map<std::string, int> obj;
  {  // .nodes
    std::ifstream f = detail::open(nodes);
    std::string line;
    std::getline(f, line);
    int header = 0;
    while (header < 2 && std::getline(f, line)) {
      if (line.rfind("NumNodes", 0) == 0 || line.rfind("NumTerminals", 0) == 0) ++header;
    }
    while (std::getline(f, line)) {
      auto t = detail::tokens(line);
      if (t.size() < 3 || t[0][0] == '#') continue;
      std::string type = t.size() > 3 ? t[3] : "";
      float w = (float)atof(t[1].c_str()), h = (float)atof(t[2].c_str());
      if (type == "terminal" || type == "terminal_NI") {
        d.term_names.push_back(t[0]);
        d.term_w.push_back(w);
        d.term_h.push_back(h);
        d.term_ni.push_back(type == "terminal_NI" ? 1 : 0);
      } else {
        d.node_names.push_back(t[0]);
        d.node_w.push_back(w);
        d.node_h.push_back(h);
      }
    }
    for (size_t i = 0; i < d.node_names.size(); ++i) obj[d.node_names[i]] = (int)i;
    for (size_t i = 0; i < d.term_names.size(); ++i) obj[d.term_names[i]] = (int)(d.node_names.size() + i);
    d.node_x.assign(d.node_w.size(), 0.f);
    d.node_y.assign(d.node_w.size(), 0.f);
    d.term_x.assign(d.term_w.size(), 0.f);
    d.term_y.assign(d.term_w.size(), 0.f);
  }
  {  // .nets
    std::ifstream f = detail::open(nets);
    std::string line;
    std::getline(f, line);
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      auto t = detail::tokens(line);
      if (t.empty() || t[0] == "NumNets" || t[0] == "NumPins") continue;
      if (t[0] != "NetDegree" || t.size() < 3) throw std::runtime_error("unsupported .nets line: " + line);
      int deg = atoi(t[2].c_str());
      for (int k = 0; k < deg; ++k) {
        if (!std::getline(f, line)) throw std::runtime_error("truncated .nets");
        auto p = detail::tokens(line);
        if (p.empty()) throw std::runtime_error("empty pin line");
        auto it = obj.find(p[0]);
        if (it == obj.end()) throw std::runtime_error("pin on unknown module " + p[0]);
        d.pin_obj.push_back(it->second);
        // "<name> <dir> : <dx> <dy>"
        d.pin_dx.push_back(p.size() > 3 ? (float)atof(p[3].c_str()) : 0.f);
        d.pin_dy.push_back(p.size() > 4 ? (float)atof(p[4].c_str()) : 0.f);
      }
      d.net_start.push_back((int64_t)d.pin_obj.size());
    }
  }
  load_pl(d, pl);
  return d;
}

// Lower-left corners, as the reference's outputBookShelf writes them.
inline void write_pl(const BookshelfDesign &d, const std::vector<float> &xy, const std::string &path) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  fprintf(f, "UCLA pl 1.0\n\n");
  for (size_t i = 0; i < d.node_names.size(); ++i)
    fprintf(f, "%s %.6f %.6f : N\n", d.node_names[i].c_str(), xy[2 * i] - 0.5f * d.node_w[i],
            xy[2 * i + 1] - 0.5f * d.node_h[i]);
  for (size_t i = 0; i < d.term_names.size(); ++i)
    fprintf(f, "%s %.6f %.6f : N %s\n", d.term_names[i].c_str(), d.term_x[i] - 0.5f * d.term_w[i],
            d.term_y[i] - 0.5f * d.term_h[i], d.term_ni[i] ? "/FIXED_NI" : "/FIXED");
  fclose(f);
}

}  // namespace eplace_amd

// deck_io.cpp -- decks from real molecules: BUDE bhff forcefields, Tripos
// mol2 molecules and the pose-grid sampler, producing exactly the deck the
// makedeck program writes (makedeck/main.cpp:49-234,
// makedeck/input-utils.cpp:72-262).  Host code only.
//
// Reproducing makedeck byte for byte fixes three things that are otherwise
// free choices: an atom's type is its entry's index within its residue
// group (not a global row), the forcefield rows are concatenated in
// residue-name order, and poses come out in the iteration order of a
// std::unordered_set keyed and hashed as makedeck keys and hashes them, with
// the six grid indices of a pose drawn left to right by the classic
// divide-and-reject mapping of 32-bit std::mt19937 words.
#include "bude_internal.h"

#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <string>
#include <tuple>
#include <unordered_set>
#include <vector>

struct bude_forcefield {
  struct Entry {
    std::string type;  // Sybyl type or atom name, as the mol2 files refer to it
    BudeFFParams params;
  };
  std::map<std::string, std::vector<Entry>> groups;  // residue name -> entries in file order
};

namespace {

struct Line {
  size_t num;        // 1-based line number in the file
  std::string text;  // without leading/trailing white space
};

bool blank(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

std::string trimmed(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && blank(s[a])) a++;
  while (b > a && blank(s[b - 1])) b--;
  return s.substr(a, b - a);
}

std::vector<std::string> columns(const std::string &s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && blank(s[i])) i++;
    size_t j = i;
    while (j < s.size() && !blank(s[j])) j++;
    if (j > i) out.push_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

// Non-blank lines of a file that are not '#' or '%' comments.
int content_lines(const char *path, const char *what, std::vector<Line> *out) {
  std::ifstream in(path);
  if (!in.good()) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string("unable to open ") + what + " file `" + path + "`");
  std::string raw;
  size_t num = 0;
  while (std::getline(in, raw)) {
    num++;
    std::string t = trimmed(raw);
    if (t.empty() || t[0] == '#' || t[0] == '%') continue;
    out->push_back(Line{num, std::move(t)});
  }
  return BUDE_OK;
}

int parse_error(const char *path, const Line &l, const std::string &what) {
  return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string("parse error at ") + path + ":" + std::to_string(l.num) +
                                                   ": " + what + " (source line: `" + l.text + "`)");
}

// A leading decimal number, as std::stof reads one (trailing text ignored).
bool to_float(const std::string &s, float *v) {
  const char *p = s.c_str();
  char *end = nullptr;
  errno = 0;
  const float x = std::strtof(p, &end);
  if (end == p || errno == ERANGE) return false;
  *v = x;
  return true;
}

bool to_count(const std::string &s, size_t *v) {
  const char *p = s.c_str();
  char *end = nullptr;
  errno = 0;
  const unsigned long long x = std::strtoull(p, &end, 10);
  if (end == p || errno == ERANGE || s[0] == '-') return false;
  *v = size_t(x);
  return true;
}

// Index in [0, n) from 32-bit engine words: the largest multiple of n that
// fits in the word range is kept, words at or above it are redrawn, and the
// kept word is divided down.  One word is consumed even for n == 1.
size_t draw_below(std::mt19937 &g, size_t n) {
  const uint64_t range = 0xffffffffull;  // engine max - min
  const uint64_t scaling = range / uint64_t(n), past = uint64_t(n) * scaling;
  uint64_t w;
  do w = uint64_t(g()); while (w >= past);
  return size_t(w / scaling);
}

using PoseKey = std::tuple<size_t, size_t, size_t, size_t, size_t, size_t>;

struct PoseKeyHash {
  size_t operator()(const PoseKey &k) const {
    return std::get<0>(k) * 100000 + std::get<1>(k) * 10000 + std::get<2>(k) * 1000 + std::get<3>(k) * 100 +
           std::get<4>(k) * 10 + std::get<5>(k);
  }
};

}  // namespace

extern "C" int bude_forcefield_read(const char *bhff_path, bude_forcefield **out) {
  if (!bhff_path || !out) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_forcefield_read: NULL argument");
  *out = nullptr;
  std::vector<Line> lines;
  if (int st = content_lines(bhff_path, "forcefield", &lines)) return st;
  auto ff = std::make_unique<bude_forcefield>();
  for (size_t r = 0; r < lines.size(); r++) {
    const std::vector<std::string> head = columns(lines[r].text);
    if (head.size() != 2) continue;  // only "RESIDUE COUNT" lines open a group
    size_t count = 0;
    if (!to_count(head[1], &count)) return parse_error(bhff_path, lines[r], "residue row needs an entry count");
    auto &group = ff->groups[head[0]];
    for (size_t b = 0; b < count; b++) {
      if (r + 1 + b >= lines.size())
        return parse_error(bhff_path, lines[r], "residue declares " + std::to_string(count) + " rows, the file ends after " +
                                                    std::to_string(b));
      const Line &l = lines[r + 1 + b];
      const std::vector<std::string> c = columns(l.text);
      if (c.size() != 9)
        return parse_error(bhff_path, l, "an atom row needs 9 columns, got " + std::to_string(c.size()));
      if (c[1].size() != 1) return parse_error(bhff_path, l, "column 2 must be a single character, got `" + c[1] + "`");
      float radius = 0, hphb = 0, scaling = 0, elsc = 0;
      if (!to_float(c[2], &radius) || !to_float(c[3], &hphb) || !to_float(c[7], &scaling) || !to_float(c[8], &elsc))
        return parse_error(bhff_path, l, "malformed number");
      BudeFFParams p;
      p.hbtype = int32_t(static_cast<signed char>(c[1][0]));
      p.radius = radius * scaling;  // the scaled radius is what the kernel uses
      p.hphb = hphb;
      p.elsc = elsc;
      group.push_back(bude_forcefield::Entry{c[0], p});
    }
  }
  if (ff->groups.empty())
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string("no residue groups in forcefield `") + bhff_path + "`");
  *out = ff.release();
  return bude::ok();
}

extern "C" int bude_forcefield_info(const bude_forcefield *ff, size_t *nresidues, size_t *nentries) {
  if (!ff) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_forcefield_info: NULL forcefield");
  size_t n = 0;
  for (const auto &g : ff->groups) n += g.second.size();
  if (nresidues) *nresidues = ff->groups.size();
  if (nentries) *nentries = n;
  return bude::ok();
}

extern "C" int bude_forcefield_params(const bude_forcefield *ff, BudeFFParams *out, size_t cap, size_t *n) {
  if (!ff || !n) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_forcefield_params: NULL argument");
  size_t total = 0;
  for (const auto &g : ff->groups) total += g.second.size();
  *n = total;
  if (!out) return bude::ok();
  if (cap < total)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_forcefield_params: room for " + std::to_string(cap) + " of " +
                                                     std::to_string(total) + " rows");
  size_t i = 0;
  for (const auto &g : ff->groups)
    for (const auto &e : g.second) out[i++] = e.params;
  return bude::ok();
}

extern "C" int bude_forcefield_destroy(bude_forcefield *ff) {
  delete ff;
  return bude::ok();
}

extern "C" int bude_mol2_read(const bude_forcefield *ff, const char *mol2_path, BudeAtom *atoms, size_t cap,
                              size_t *natoms, size_t *nconformations) {
  if (!ff || !mol2_path || !natoms) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mol2_read: NULL argument");
  std::vector<Line> lines;
  if (int st = content_lines(mol2_path, "mol2", &lines)) return st;
  static const std::string kBegin = "@<TRIPOS>ATOM", kSection = "@<TRIPOS>";
  size_t begin = 0;
  while (begin < lines.size() && lines[begin].text != kBegin) begin++;
  if (begin == lines.size())
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string("`") + mol2_path + "`: atom section marker (" + kBegin +
                                                     ") missing");
  begin++;
  size_t end = begin;
  while (end < lines.size() && lines[end].text.compare(0, kSection.size(), kSection) != 0) end++;
  if (end == lines.size())
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string("`") + mol2_path + "`: no section (" + kSection +
                                                     "...) after the atoms");

  // one residue group: atoms are looked up by Sybyl type; several: by
  // residue and atom name
  const bool by_type = ff->groups.size() == 1;
  const size_t want_cols = by_type ? 9 : 10;
  std::vector<BudeAtom> out;
  for (size_t i = begin; i < end; i++) {
    const Line &l = lines[i];
    const std::vector<std::string> c = columns(l.text);
    if (c.size() != want_cols)
      return parse_error(mol2_path, l,
                         std::string("a forcefield with ") + (by_type ? "one residue group" : "several residue groups") +
                             " needs " + std::to_string(want_cols) + "-column atom rows, got " + std::to_string(c.size()));
    if (c[5] == "H" || c[5] == "h") continue;  // hydrogens are not scored
    const std::string &type = by_type ? c[5] : c[1];
    const std::string residue = (by_type ? ff->groups.begin()->first : c[7]).substr(0, 3);
    BudeAtom a;
    if (!to_float(c[2], &a.x) || !to_float(c[3], &a.y) || !to_float(c[4], &a.z))
      return parse_error(mol2_path, l, "malformed coordinate");
    const auto g = ff->groups.find(residue);
    size_t idx = 0;
    if (g != ff->groups.end())
      while (idx < g->second.size() && g->second[idx].type != type) idx++;
    if (g == ff->groups.end() || idx == g->second.size())
      return parse_error(mol2_path, l, "no forcefield entry " + residue + "." + type);
    a.type = int32_t(idx);
    out.push_back(a);
  }

  // "@<BUDE>CONF n" blocks: one x y z row per heavy atom
  size_t nconf = 0;
  for (size_t i = end; i < lines.size(); i++) {
    const std::vector<std::string> c = columns(lines[i].text);
    if (c.size() != 2 || c[0] != "@<BUDE>CONF") continue;
    for (size_t a = 0; a < out.size(); a++) {
      if (i + 1 + a >= lines.size()) return parse_error(mol2_path, lines[i], "conformation ends early");
      const Line &l = lines[i + 1 + a];
      const std::vector<std::string> xyz = columns(l.text);
      float v = 0;
      if (xyz.size() != 3 || !to_float(xyz[0], &v) || !to_float(xyz[1], &v) || !to_float(xyz[2], &v))
        return parse_error(mol2_path, l, "a conformation row needs 3 numbers");
    }
    nconf++;
  }

  *natoms = out.size();
  if (nconformations) *nconformations = nconf;
  if (!atoms) return bude::ok();
  if (cap < out.size())
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mol2_read: room for " + std::to_string(cap) + " of " +
                                                     std::to_string(out.size()) + " atoms");
  std::copy(out.begin(), out.end(), atoms);
  return bude::ok();
}

extern "C" int bude_poses_generate(size_t nposes, uint64_t seed, const float *const values[6], const size_t counts[6],
                                   float *const rows[6]) {
  if (!values || !counts || !rows) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_poses_generate: NULL argument");
  size_t combos = 1;
  for (int k = 0; k < 6; k++) {
    if (counts[k] == 0 || !values[k] || (nposes && !rows[k]))
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_poses_generate: every field needs values and an output row");
    if (counts[k] > 0xffffffffull)
      return bude::fail(BUDE_ERR_UNSUPPORTED, "bude_poses_generate: more than 2^32 values in one field");
    combos = combos > std::numeric_limits<size_t>::max() / counts[k] ? std::numeric_limits<size_t>::max()
                                                                      : combos * counts[k];
  }
  if (nposes > combos)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_poses_generate: " + std::to_string(nposes) +
                          " poses exceed the grid's " + std::to_string(combos) + " combinations");
  std::unordered_set<PoseKey, PoseKeyHash> set(nposes);
  std::mt19937 g(static_cast<std::mt19937::result_type>(seed));
  while (set.size() < nposes) {
    // one statement per field: the draw order is part of the deck
    const size_t t = draw_below(g, counts[0]);
    const size_t r = draw_below(g, counts[1]);
    const size_t p = draw_below(g, counts[2]);
    const size_t x = draw_below(g, counts[3]);
    const size_t y = draw_below(g, counts[4]);
    const size_t z = draw_below(g, counts[5]);
    set.emplace(t, r, p, x, y, z);
  }
  constexpr double kDegToRad = M_PI / 180.0;
  size_t i = 0;
  for (const PoseKey &k : set) {
    rows[0][i] = float(double(values[0][std::get<0>(k)]) * kDegToRad);
    rows[1][i] = float(double(values[1][std::get<1>(k)]) * kDegToRad);
    rows[2][i] = float(double(values[2][std::get<2>(k)]) * kDegToRad);
    rows[3][i] = values[3][std::get<3>(k)];
    rows[4][i] = values[4][std::get<4>(k)];
    rows[5][i] = values[5][std::get<5>(k)];
    i++;
  }
  return bude::ok();
}

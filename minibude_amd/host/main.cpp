// main.cpp -- bude-mi355x, the driver of the mi355x model.
//
// Same command line, deck handling, validation, accounting and report as the
// reference driver (src/main.cpp), so scripts written for `<model>-bude`
// (including heatmap.py over --csv) run unchanged:
//   -d/--device, -i/--iter, -n/--poses, -p/--ppwi (CSV or `all`),
//   -w/--wgsize (CSV), --deck DIR, -o/--out PATH, -r/--rows N, --csv,
//   -l/--list, -h/--help
// plus --out-full PATH (every energy at %.9g, which -o's 2 significant digits
// do not give).
//
// Behaviours kept on purpose (src/main.cpp line numbers):
//   * PPWI is one of 1,2,4,8,16,32,64,128 (:48, :221-231); default ppwi 1 and
//     wgsize 1 (:282-283) -- wgsize 1 is not a whole wavefront, so the model
//     reports it incompatible: the combination is skipped with the library's
//     reason before any device buffer is made, and the exit status is failure;
//   * combos with nposes < ppwi*wgsize or nposes % (ppwi*wgsize) != 0 are
//     skipped with a warning (:520-530);
//   * validation: signed |ref - e| / ref, both |.| < 1 skipped, fail above
//     0.025 % (:329-342, :371); an invalid result reports DBL_MAX ms (:390-392);
//   * FLOP / FInst / interaction accounting (:358-369);
//   * best = min sum_ms (:550-559); exit status = all results valid (:561).
// An empty result set (every combo skipped) prints "best: ~" instead of
// dereferencing an empty range.
#include "bude.h"
#include "mi355x/fasten.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr size_t kPpwis[] = {1, 2, 4, 8, 16, 32, 64, 128};

template <typename C> std::string join(const C &xs, const std::string &sep = ",") {
  std::ostringstream o;
  bool first = true;
  for (const auto &x : xs) {
    if (!first) o << sep;
    o << x;
    first = false;
  }
  return o.str();
}

struct Stats {  // over the non-warmup iterations
  double min = 0, max = 0, sum = 0, mean = 0, stdDev = 0;
  static Stats of(const std::vector<double> &ms) {
    Stats s;
    if (ms.empty()) return s;
    s.min = *std::min_element(ms.begin(), ms.end());
    s.max = *std::max_element(ms.begin(), ms.end());
    s.sum = std::accumulate(ms.begin(), ms.end(), 0.0);
    s.mean = s.sum / double(ms.size());
    double v = 0;
    for (double x : ms) v += (x - s.mean) * (x - s.mean);
    s.stdDev = std::sqrt(v / double(ms.size()));
    return s;
  }
  static Stats invalid() {
    Stats s;
    s.min = s.max = s.sum = s.mean = DBL_MAX;
    return s;
  }
};

struct Result {
  bool valid;
  double maxDiffPct;
  Sample sample;
  Stats ms;
  double gflops, ginsts, interactionsPerSec;
  std::vector<size_t> failed;
};

[[noreturn]] void die(const std::string &msg) {
  std::cerr << msg << std::endl;
  std::exit(EXIT_FAILURE);
}

template <typename T> std::vector<T> readRecords(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  if (!in.good()) throw std::invalid_argument("Bad file: " + path);
  std::vector<char> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  std::vector<T> xs(bytes.size() / sizeof(T));
  std::copy_n(bytes.data(), xs.size() * sizeof(T), reinterpret_cast<char *>(xs.data()));
  return xs;
}

size_t parseCount(const std::string &s, const std::string &what) {
  try {
    size_t used = 0;
    const long v = std::stol(s, &used);
    if (v < 0) die("positive integer required for <" + what + ">: `" + s + "`");
    return size_t(v);
  } catch (const std::exception &) {
    die("malformed value, integer required for <" + what + ">: `" + s + "`");
  }
}

std::vector<size_t> parseCounts(const std::string &s, const std::string &what) {
  std::vector<size_t> out;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, ',')) out.push_back(parseCount(tok, what));
  return out;
}

void usage() {
  std::cout << "\nUsage: ./bude-mi355x [COMMAND|OPTIONS]\n\n"
            << "Commands:\n"
            << "  help -h --help       Print this message\n"
            << "  list -l --list       List available devices\n"
            << "Options:\n"
            << "  -d --device  INDEX   Select device at INDEX from output of --list, performs a substring match of "
               "device names if INDEX is not an integer\n"
            << "                       [optional] default=0\n"
            << "  -i --iter    I       Repeat kernel I times\n"
            << "                       [optional] default=" << DEFAULT_ITERS << "\n"
            << "  -n --poses   N       Compute energies for only N poses, use 0 for deck max\n"
            << "                       [optional] default=0 \n"
            << "  -p --ppwi    PPWI    A CSV list of poses per work-item for the kernel, use `all` for everything\n"
            << "                       [optional] default=1; available=" << join(kPpwis) << "\n"
            << "  -w --wgsize  WGSIZE  A CSV list of work-group sizes (64,128,256,512,1024 on this model)\n"
            << "                       [optional] default=1\n"
            << "     --deck    DIR     Use the DIR directory as input deck\n"
            << "                       [optional] default=`" << DATA_DIR << "`\n"
            << "  -o --out     PATH    Save resulting energies to PATH (2 significant digits, as the reference)\n"
            << "                       [optional]\n"
            << "     --out-full PATH   Save resulting energies to PATH at full precision (%.9g)\n"
            << "                       [optional]\n"
            << "  -r --rows    N       Output first N row(s) of energy values as part of the on-screen result\n"
            << "                       [optional] default=" << DEFAULT_ENERGY_ENTRIES << "\n"
            << "     --csv             Output results in CSV format\n"
            << "                       [optional] default=false" << std::endl;
}

struct Cli {
  Params params;
  std::vector<size_t> wgsizes, ppwis;
  std::string outFull;
};

Cli parse(const std::vector<std::string> &args) {
  Cli c;
  Params &p = c.params;
  p.deckDir = DATA_DIR;
  size_t nposes = 0;
  auto value = [&](size_t &i, const std::string &flag) -> const std::string & {
    if (i + 1 >= args.size()) die("[" + flag + "] specified but no value was given");
    return args[++i];
  };
  for (size_t i = 0; i < args.size(); ++i) {
    const std::string &a = args[i];
    if (a == "-i" || a == "--iter") p.iterations = parseCount(value(i, a), "iter");
    else if (a == "-n" || a == "--poses") nposes = parseCount(value(i, a), "poses");
    else if (a == "-d" || a == "--device") p.deviceSelector = value(i, a);
    else if (a == "--deck") p.deckDir = value(i, a);
    else if (a == "-o" || a == "--out") p.output = value(i, a);
    else if (a == "--out-full") c.outFull = value(i, a);
    else if (a == "-r" || a == "--rows") p.outRows = parseCount(value(i, a), "rows");
    else if (a == "-w" || a == "--wgsize") c.wgsizes = parseCounts(value(i, a), "wgsize");
    else if (a == "-p" || a == "--ppwi") {
      const std::string &v = value(i, a);
      if (v == "all") c.ppwis.assign(std::begin(kPpwis), std::end(kPpwis));
      else {
        c.ppwis = parseCounts(v, "ppwi");
        for (size_t q : c.ppwis)
          if (std::find(std::begin(kPpwis), std::end(kPpwis), q) == std::end(kPpwis))
            die("PPWI " + std::to_string(q) + " is not a supported value, should be one of `" + join(kPpwis) + "`");
      }
    } else if (a == "--csv") p.csv = true;
    else if (a == "list" || a == "--list" || a == "-l") p.list = true;
    else if (a == "help" || a == "--help" || a == "-h") {
      usage();
      std::exit(EXIT_SUCCESS);
    } else {
      std::cout << "Unrecognized argument '" << a << "' (try '--help')" << std::endl;
      std::exit(EXIT_FAILURE);
    }
  }
  if (c.ppwis.empty()) c.ppwis = {kPpwis[0]};
  if (c.wgsizes.empty()) c.wgsizes = {1};
  if (p.list) return c;

  p.ligand = readRecords<Atom>(p.deckDir + FILE_LIGAND);
  p.protein = readRecords<Atom>(p.deckDir + FILE_PROTEIN);
  p.forcefield = readRecords<FFParams>(p.deckDir + FILE_FORCEFIELD);
  const std::vector<float> poses = readRecords<float>(p.deckDir + FILE_POSES);
  if (poses.size() % 6 != 0)
    throw std::invalid_argument("Pose size (" + std::to_string(poses.size()) + ") not divisible my 6!");
  p.maxPoses = poses.size() / 6;
  if (nposes == 0) nposes = p.maxPoses;
  if (nposes > p.maxPoses)
    throw std::invalid_argument("Requested poses (" + std::to_string(nposes) + ") exceeded max poses (" +
                                std::to_string(p.maxPoses) + ") for deck");
  for (size_t k = 0; k < 6; ++k)
    p.poses[k].assign(poses.begin() + long(k * p.maxPoses), poses.begin() + long(k * p.maxPoses + nposes));

  std::ifstream ref(p.deckDir + FILE_REF_ENERGIES);
  for (std::string line; std::getline(ref, line);) p.refEnergies.push_back(std::stof(line));
  if (p.nposes() > p.refEnergies.size())
    throw std::invalid_argument("Size of reference energies (" + std::to_string(p.refEnergies.size()) +
                                ") is less than poses (" + std::to_string(p.nposes()) + ")");
  return c;
}

// Signed relative difference of the reference driver (src/main.cpp:329).
double difference(float ref, float e) { return std::fabs(double(ref) - e) / ref; }

Result evaluate(const Params &p, Sample s) {
  std::vector<double> ms;
  for (size_t i = p.warmupIterations; i < s.kernelTimes.size(); ++i)
    ms.push_back(elapsedMillis(s.kernelTimes[i].first, s.kernelTimes[i].second));
  const double runtime = (std::accumulate(ms.begin(), ms.end(), 0.0) / double(p.iterations)) * 1e-3;

  const double P = double(s.ppwi), L = double(p.natlig()), N = double(p.natpro()), n = double(p.nposes());
  const double opsPerGroup = P * 27 + L * (2 + P * 18 + N * (10 + P * 30)) + P;
  const double flops = opsPerGroup * (n / P);
  const double finsts = 25.0 * N * L * n;
  const double interactions = n * L * N;

  double maxdiff = 0.0;
  std::vector<size_t> failed;
  for (size_t i = 0; i < s.energies.size(); ++i) {
    if (std::fabs(p.refEnergies[i]) < 1.f && std::fabs(s.energies[i]) < 1.f) continue;
    const double d = difference(p.refEnergies[i], s.energies[i]);
    maxdiff = std::max(maxdiff, d);
    if (100.0 * d > DIFF_TOLERANCE_PCT) failed.push_back(i);
  }
  const double maxDiffPct = 100.0 * maxdiff;
  const bool valid = maxDiffPct < DIFF_TOLERANCE_PCT;
  if (!valid) {
    std::cerr << "# Verification failed for ppwi=" << s.ppwi << ", wgsize=" << s.wgsize
              << "; difference exceeded tolerance (" << DIFF_TOLERANCE_PCT << "%)\n"
              << "# Bad energies (failed/total=" << failed.size() << "/" << s.energies.size() << ", showing first "
              << std::min(failed.size(), p.outRows) << "): \n# index,actual,expected,difference_%\n";
    for (size_t k = 0; k < std::min(failed.size(), p.outRows); ++k) {
      const size_t i = failed[k];
      std::cerr << "# " << i << "," << s.energies[i] << "," << p.refEnergies[i] << ","
                << difference(p.refEnergies[i], s.energies[i]) * 100.0 << "\n";
    }
    std::cerr << std::flush;
    return {false, maxDiffPct, std::move(s), Stats::invalid(), 0, 0, 0, failed};
  }
  return {true, maxDiffPct, std::move(s), Stats::of(ms), flops / runtime / 1e9, finsts / runtime / 1e9,
          interactions / runtime, failed};
}

void dumpEnergies(const Params &p, const std::string &outFull, const Result &r) {
  if (!p.output.empty()) {
    std::ofstream out(p.output, std::ios::trunc);
    for (float e : r.sample.energies) out << std::setw(7) << std::setprecision(2) << e << "\n";
  }
  if (!outFull.empty()) {
    std::ofstream out(outFull, std::ios::trunc);
    char buf[32];
    for (float e : r.sample.energies) {
      std::snprintf(buf, sizeof(buf), "%.9g\n", double(e));
      out << buf;
    }
  }
}

void showYaml(const Params &p, const Result &r) {
  std::vector<double> raw;
  for (const auto &t : r.sample.kernelTimes) raw.push_back(elapsedMillis(t.first, t.second));
  const std::string contextMs =
      r.sample.contextTime ? std::to_string(elapsedMillis(r.sample.contextTime->first, r.sample.contextTime->second))
                           : "~";
  std::cout.precision(3);
  std::cout << std::fixed << "  - outcome:             { valid: " << (r.valid ? "true" : "false")
            << ", max_diff_%: " << r.maxDiffPct << " }\n"
            << "    param:               { ppwi: " << r.sample.ppwi << ", wgsize: " << r.sample.wgsize << " }\n"
            << "    raw_iterations:      [" << join(raw) << "]\n"
            << "    context_ms:          " << contextMs << "\n"
            << "    sum_ms:              " << r.ms.sum << "\n"
            << "    avg_ms:              " << r.ms.mean << "\n"
            << "    min_ms:              " << r.ms.min << "\n"
            << "    max_ms:              " << r.ms.max << "\n"
            << "    stddev_ms:           " << r.ms.stdDev << "\n"
            << "    giga_interactions/s: " << r.interactionsPerSec / 1e9 << "\n"
            << "    gflop/s:             " << r.gflops << "\n"
            << "    gfinst/s:            " << r.ginsts << "\n"
            << "    energies:            \n";
  for (size_t i = 0; i < std::min(p.outRows, r.sample.energies.size()); ++i)
    std::cout << std::setprecision(2) << "      - " << r.sample.energies[i] << "\n";
  std::cout << std::flush;
}

void showCsv(const Result &r, bool header) {
  if (header) std::cout << "ppwi,wgsize,sum_ms,avg_ms,min_ms,max_ms,stddev_ms,interactions/s,gflops/s,gfinst/s\n";
  std::cout.precision(3);
  std::cout << std::fixed << r.sample.ppwi << "," << r.sample.wgsize << "," << r.ms.sum << "," << r.ms.mean << ","
            << r.ms.min << "," << r.ms.max << "," << r.ms.stdDev << "," << r.interactionsPerSec << "," << r.gflops
            << "," << r.ginsts << std::endl;
}

// Device selection: index, else substring of the name, else the only device
// (src/main.cpp:398-419).
std::pair<long, std::string> selectDevice(const std::string &needle, const std::vector<Device> &devices) {
  if (devices.empty()) return {-1, ""};
  if (needle.empty()) return {long(devices[0].first), devices[0].second};
  try {
    size_t used = 0;
    const unsigned long idx = std::stoul(needle, &used);
    if (used == needle.size() && idx < devices.size()) return {long(devices[idx].first), devices[idx].second};
    throw std::out_of_range("device index " + needle);
  } catch (const std::exception &e) {
    std::cerr << "# Unable to parse/select device index `" << needle << "`:" << e.what() << std::endl;
    std::cerr << "# Attempting to match device with substring  `" << needle << "`" << std::endl;
    for (const auto &d : devices)
      if (d.second.find(needle) != std::string::npos) return {long(d.first), d.second};
    if (devices.size() == 1) {
      std::cerr << "# No matching device but there's only one device, using it anyway" << std::endl;
      return {long(devices[0].first), devices[0].second};
    }
    std::cerr << "# No matching devices" << std::endl;
    return {-1, ""};
  }
}

using KernelFn = std::function<Sample(const Params &, size_t, size_t)>;

template <size_t N> void registerPpwi(std::map<size_t, KernelFn> &k) {
  k[N] = [](const Params &p, size_t wgsize, size_t device) {
    Mi355xBude<N> model;
    const auto copy = std::make_unique<Params>(p);
    return model.fasten(*copy, wgsize, device);
  };
}

// Bude<PPWI>::compatible for a runtime ppwi; the reason goes to `why`.
// The reference driver only warns here (src/main.cpp:492-495) and calls
// fasten anyway; this driver skips the combination instead, before any
// device buffer is made, since the library cannot run it.
bool compatible(size_t ppwi, size_t wgsize, size_t device, std::string &why) {
  int ok = 0;
  if (bude_mi355x_compatible(int(device), ppwi, wgsize, &ok) == BUDE_OK && ok == 1) return true;
  why = bude_mi355x_last_error();
  return false;
}

bool run(const Cli &c) {
  const Params &p = c.params;
  std::map<size_t, KernelFn> kernel;
  registerPpwi<1>(kernel);
  registerPpwi<2>(kernel);
  registerPpwi<4>(kernel);
  registerPpwi<8>(kernel);
  registerPpwi<16>(kernel);
  registerPpwi<32>(kernel);
  registerPpwi<64>(kernel);
  registerPpwi<128>(kernel);

  const std::vector<Device> devices = Mi355xBude<1>().enumerateDevices();
  if (devices.empty()) std::cerr << " # (no devices available)" << std::endl;
  if (p.list) {
    std::cout << (p.csv ? "index,name" : "devices:") << std::endl;
    for (size_t j = 0; j < devices.size(); ++j) {
      if (p.csv) std::cout << j << "," << devices[j].second << std::endl;
      else std::cout << "  " << j << ": \"" << devices[j].second << "\"" << std::endl;
    }
    return true;
  }
  const auto dev = selectDevice(p.deviceSelector, devices);
  if (dev.first < 0) return false;
  if (!p.csv) std::cout << "device: { index: " << dev.first << ", " << " name: \"" << dev.second << "\" }" << std::endl;

  std::vector<Result> results;
  bool dumped = false, skippedIncompatible = false;
  for (size_t ppwi : c.ppwis) {
    for (size_t wgsize : c.wgsizes) {
      std::string why;
      if (!compatible(ppwi, wgsize, size_t(dev.first), why)) {
        std::cerr << "Selected device is not compatible with this implementation (ppwi " << ppwi << ", wgsize "
                  << wgsize << ": " << why << "), skipping" << std::endl;
        skippedIncompatible = true;
        continue;
      }
      if (p.nposes() < ppwi * wgsize) {
        std::cout << " # WARNING: pose count " << p.nposes() << " <= (" << wgsize << " (wgsize) * " << ppwi
                  << " (ppwi)), skipping" << std::endl;
        continue;
      }
      if (p.nposes() % (ppwi * wgsize) != 0) {
        std::cout << " # WARNING: pose count " << p.nposes() << " % (" << wgsize << " (wgsize) * " << ppwi
                  << " (ppwi)) != 0, skipping" << std::endl;
        continue;
      }
      results.push_back(evaluate(p, kernel.at(ppwi)(p, wgsize, size_t(dev.first))));
      const Result &r = results.back();
      std::cout << "# (ppwi=" << ppwi << ",wgsize=" << wgsize << ",valid=" << r.valid << ")" << std::endl;
      if (!dumped || !r.valid) {
        dumped = true;
        dumpEnergies(p, c.outFull, r);
      }
    }
  }

  if (!p.csv) std::cout << "results:" << std::endl;
  for (size_t i = 0; i < results.size(); ++i) {
    if (p.csv) showCsv(results[i], i == 0);
    else showYaml(p, results[i]);
  }
  if (results.empty()) {
    std::cout << (p.csv ? "# " : "") << "best: ~" << std::endl;
    return !skippedIncompatible;
  }
  const auto best = std::min_element(results.begin(), results.end(),
                                     [](const Result &a, const Result &b) { return a.ms.sum < b.ms.sum; });
  std::cout << (p.csv ? "# " : "") << "best: { min_ms: " << best->ms.min << ", max_ms: " << best->ms.max
            << ", sum_ms: " << best->ms.sum << ", avg_ms: " << best->ms.mean << ", ppwi: " << best->sample.ppwi
            << ", wgsize: " << best->sample.wgsize << " }\n";
  return !skippedIncompatible &&
         std::all_of(results.begin(), results.end(), [](const Result &r) { return r.valid; });
}

}  // namespace

int main(int argc, char *argv[]) {
  const Cli c = parse(std::vector<std::string>(argv + 1, argv + argc));
  const Params &p = c.params;
  if (!p.csv) {
    const std::time_t t = std::time(nullptr);
    std::cout << "miniBUDE:  mi355x model (libbude_mi355x ABI " << bude_mi355x_abi_version() << ")\n"
              << "compile_commands:\n   - \"hipcc --offload-arch=gfx950 -O3 (see minibude_amd/build.py)\"\n"
              << "vcs:\n   # ~\nhost_cpu:\n  ~\n"
              << "time: { epoch_s:" << t << ", formatted: \"" << std::put_time(std::gmtime(&t), "%c %Z") << "\" }"
              << std::endl;
    if (!p.list) {
      std::cout << "deck:\n"
                << "  path:         \"" << p.deckDir << "\"\n"
                << "  poses:        " << p.maxPoses << "\n"
                << "  proteins:     " << p.natpro() << "\n"
                << "  ligands:      " << p.natlig() << "\n"
                << "  forcefields:  " << p.ntypes() << "\n"
                << "config:\n"
                << "  iterations:   " << p.iterations << "\n"
                << "  poses:        " << p.nposes() << "\n"
                << "  ppwi:\n"
                << "    available:  [" << join(kPpwis) << "]\n"
                << "    selected:   [" << join(c.ppwis) << "]\n"
                << "  wgsize:       [" << join(c.wgsizes) << "]" << std::endl;
    }
  }
  try {
    return run(c) ? EXIT_SUCCESS : EXIT_FAILURE;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
}

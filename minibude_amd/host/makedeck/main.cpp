// main.cpp -- bude-makedeck: a deck for the benchmark from a BUDE bhff
// forcefield and protein/ligand mol2 files, with makedeck's command line and
// output (makedeck/main.cpp:49-234):
//   -f/--forcefield BHFF  -p/--protein MOL2  -l/--ligand MOL2
//   -s/--pose-seed N (42)  -n/--pose-length N (65536)  -o/--out DIR  --force
//   -h/--help
// The deck directory gets forcefield.in, protein.in, ligand.in, poses.in
// (byte-identical to makedeck's for the same inputs), ref_energies.out
// (one energy per line, fixed with 5 decimals) and params.txt (the
// arguments).  makedeck scored the poses with its CPU reference kernel; here
// the library's plain kernel scores them on GPU `--device` (default 0), and
// without a usable device the program stops with the library's message.
//
// Kept from makedeck on purpose: the pose count must be a multiple of 256
// (its kernel's work-group size, makedeck/main.cpp:123-125); an existing
// output path is an error unless --force, which deletes it first.
#include "bude_mi355x.h"

#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

namespace fs = std::filesystem;

namespace {

[[noreturn]] void die(const std::string &why) {
  std::cerr << why << std::endl;
  std::exit(EXIT_FAILURE);
}

void check(int status, const std::string &what) {
  if (status != BUDE_OK) die(what + ": " + bude_mi355x_last_error());
}

size_t count_arg(const std::string &v, const std::string &name) {
  char *end = nullptr;
  const long long x = std::strtoll(v.c_str(), &end, 10);
  if (end == v.c_str() || *end) die("malformed value, integer required for <" + name + ">: `" + v + "`");
  if (x < 0) die("positive integer required for " + name + ": `" + v + "`");
  return size_t(x);
}

void usage() {
  std::cout << "Builds an input deck for the benchmark from BUDE forcefield and molecule files.\n"
               "Usage: bude-makedeck [OPTIONS]\n\n"
               "Options:\n"
               "  -h  --help         Print this message\n"
               "  -f  --forcefield   Path to a forcefield bhff file\n"
               "  -p  --protein      Path to a protein mol2 file\n"
               "  -l  --ligand       Path to a ligand mol2 file\n"
               "  -s  --pose-seed    The random seed used to generate pose combinations (default: 42)\n"
               "  -n  --pose-length  The amount of poses to generate (default: 65536)\n"
               "  -o  --out          The output directory (gets {protein,ligand,forcefield,poses}.in,\n"
               "                     ref_energies.out and params.txt)\n"
               "      --force        Delete an existing file or directory at the output path first\n"
               "  -d  --device       GPU that scores the poses for ref_energies.out (default: 0)\n"
            << std::endl;
}

template <typename T>
void write_raw(const fs::path &p, const T *data, size_t n) {
  std::ofstream out(p, std::ios::binary | std::ios::app);
  out.write(reinterpret_cast<const char *>(data), std::streamsize(n * sizeof(T)));
  if (!out) die("unable to write " + p.string());
}

}  // namespace

int main(int argc, char **argv) {
  const std::vector<std::string> args(argv + 1, argv + argc);
  std::string bhff, protein_mol2, ligand_mol2, out_dir;
  size_t seed = 42, nposes = 65536, device = 0;
  bool force = false;

  for (size_t i = 0; i < args.size(); ++i) {
    const std::string &a = args[i];
    auto value = [&](const char *names) -> std::string {
      if (i + 1 >= args.size()) die(std::string("[") + names + "] specified but no value was given");
      return args[++i];
    };
    auto path = [&](const char *names, const char *what) {
      const std::string v = value(names);
      if (!std::ifstream(v).good()) die(std::string("Unable to open ") + what + " file: `" + v + "`, is the path valid?");
      return v;
    };
    if (a == "-f" || a == "--forcefield") bhff = path("--forcefield-f", "forceFieldBhff");
    else if (a == "-p" || a == "--protein") protein_mol2 = path("--protein-p", "proteinMol2");
    else if (a == "-l" || a == "--ligand") ligand_mol2 = path("--ligand-l", "ligandMol2");
    else if (a == "-s" || a == "--pose-seed") seed = count_arg(value("--pose-seed-s"), "poseSeed");
    else if (a == "-n" || a == "--pose-length") nposes = count_arg(value("--pose-length-n"), "nPoses");
    else if (a == "-o" || a == "--out") out_dir = value("--out-o");
    else if (a == "-d" || a == "--device") device = count_arg(value("--device-d"), "device");
    else if (a == "--force") force = true;
    else if (a == "-h" || a == "--help") {
      usage();
      return EXIT_SUCCESS;
    } else die("Unrecognized argument '" + a + "' (try '--help')");
  }

  if (nposes % 256 != 0)
    die("poseSize(" + std::to_string(nposes) + ") cannot be divided by WGSIZE(256)");
  if (bhff.empty()) die("--forcefield is a required parameter");
  if (protein_mol2.empty()) die("--protein is a required parameter");
  if (ligand_mol2.empty()) die("--ligand is a required parameter");
  if (out_dir.empty()) die("--out is a required parameter");

  const fs::path dir(out_dir);
  std::error_code ec;
  if (fs::is_regular_file(dir, ec) || fs::is_directory(dir, ec)) {
    if (!force) die("Output path already exists (either a file or directory), pass --force to delete/overwrite them");
    fs::remove_all(dir, ec);
    if (ec) die("Unable to delete " + dir.string() + ": " + ec.message());
  }
  if (!fs::create_directories(dir, ec)) die("Unable to create output directory: " + dir.string());

  bude_forcefield *ff = nullptr;
  check(bude_forcefield_read(bhff.c_str(), &ff), "forcefield");
  size_t nres = 0, nff = 0;
  check(bude_forcefield_info(ff, &nres, &nff), "forcefield");
  if (nres != 1)
    std::cout << "The forcefield has " << nres << " residue groups: their rows are concatenated in residue-name "
                 "order and a mol2 atom's type is its entry's index within its residue, as makedeck writes them."
              << std::endl;
  std::vector<BudeFFParams> params(nff);
  check(bude_forcefield_params(ff, params.data(), params.size(), &nff), "forcefield");

  auto molecule = [&](const std::string &mol2, size_t *nconf) {
    size_t n = 0;
    check(bude_mol2_read(ff, mol2.c_str(), nullptr, 0, &n, nconf), mol2);
    std::vector<BudeAtom> atoms(n);
    check(bude_mol2_read(ff, mol2.c_str(), atoms.data(), atoms.size(), &n, nconf), mol2);
    return atoms;
  };
  size_t pconf = 0, lconf = 0;
  const std::vector<BudeAtom> protein = molecule(protein_mol2, &pconf);
  const std::vector<BudeAtom> ligand = molecule(ligand_mol2, &lconf);
  check(bude_forcefield_destroy(ff), "forcefield");

  // makedeck's grid (makedeck/main.cpp:12-20): 10 degree steps, 1 A steps
  std::vector<float> degrees, trans;
  for (int d = -170; d <= 180; d += 10) degrees.push_back(float(d));
  for (int t = -7; t <= 7; ++t) trans.push_back(float(t));
  const float *values[6] = {degrees.data(), degrees.data(), degrees.data(), trans.data(), trans.data(), trans.data()};
  const size_t counts[6] = {degrees.size(), degrees.size(), degrees.size(), trans.size(), trans.size(), trans.size()};
  std::vector<float> poses(6 * nposes);
  float *rows[6];
  for (int k = 0; k < 6; k++) rows[k] = poses.data() + size_t(k) * nposes;
  check(bude_poses_generate(nposes, seed, values, counts, rows), "poses");

  write_raw(dir / "forcefield.in", params.data(), params.size());
  write_raw(dir / "protein.in", protein.data(), protein.size());
  write_raw(dir / "ligand.in", ligand.data(), ligand.size());
  write_raw(dir / "poses.in", poses.data(), poses.size());

  std::vector<float> energies(nposes);
  double ms = 0;
  if (nposes) {
    std::cout << "Scoring " << nposes << " poses on device " << device << " ..." << std::endl;
    const float *prow[6] = {rows[0], rows[1], rows[2], rows[3], rows[4], rows[5]};
    bude_ctx *ctx = nullptr;
    check(bude_mi355x_create(int(device), protein.data(), protein.size(), ligand.data(), ligand.size(),
                             params.data(), params.size(), prow, nposes, &ctx),
          "device context");
    check(bude_mi355x_fasten(ctx, 1, 256, 0, 1, energies.data(), nullptr, &ms), "fasten");
    check(bude_mi355x_destroy(ctx), "device context");
  }

  {
    std::ofstream ref(dir / "ref_energies.out");
    ref << std::fixed << std::setprecision(5);
    for (float e : energies) ref << e << "\n";
    if (!ref) die("unable to write ref_energies.out");
  }
  {
    std::ofstream p(dir / "params.txt");
    for (const auto &a : args) p << a << " ";
    p << std::endl;
  }

  std::cout << "Output:\n"
            << "- Forcefield: " << nres << " residue group(s), " << params.size() << " row(s)\n"
            << "- Protein: " << protein.size() << " atom(s), " << pconf << " conformation(s)\n"
            << "- Ligand: " << ligand.size() << " atom(s), " << lconf << " conformation(s)\n"
            << "- Poses: " << nposes << " (seed=" << seed << ")\n"
            << "- Reference energy: " << energies.size() << " lines (kernel: " << std::setprecision(3) << ms
            << " ms)\n"
            << "Done" << std::endl;
  return EXIT_SUCCESS;
}

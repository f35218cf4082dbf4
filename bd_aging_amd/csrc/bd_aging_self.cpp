// bd_aging_self -- drop-in driver of the aging workflow (main.cu:56-1037) on
// the MI355X engine (libbd_hip.so, include/bd_abi.h).
//
//   bd_aging_self < bd_aging_self.data
//       the reference's 11-line stdin protocol (main.cu:93-148; the file
//       wca_aging_linear.data is read first when present, main.cu:90-91),
//       ../configs/[decadeK/]init_config_i inputs and results/[decadeK/]*.out
//       outputs (main.cu:885-980).
//   bd_aging_self --aging < bd_aging_self.data
//       the same run; with log sampling and decade dumps it also measures F_s
//       and the MSD from every dumped waiting time inside this one run and
//       writes them to results/decadeK/ (lag, value), in place of one
//       reheat-and-restart run per decade (run-fself.sh:149-197).
//   bd_aging_self --system < bd_aging.data
//       the effective-force workflow of run-system.sh:118-145: aging run plus
//       the Contracting_Vector3D force contraction, results/force_contraction.out.
//   bd_aging_self --make-configs NBIG NSML PHI SIGMA_BIG SIGMA_SML NCONFIGS SEED DIR
//                 [T EQ_STEPS DT]
//       writes DIR/init_config_1..NCONFIGS: random placement, overlap removal
//       on the GPU (harmonic contact potential at T=0) and an optional
//       equilibration at temperature T (the role of the external
//       set_init_config / eliminate_overlap / bd_aging_configs binaries that
//       run-fself.sh:131-138 calls and the reference does not contain).
//
// Differences from main.cu, each a fix of a reference defect:
//   * the n_configs ensemble runs as concurrent replicas of one engine, each
//     with its own rannew64 seed (the reference reuses idum for every config,
//     main.cu:483-484);
//   * log sampling needs no compile flag (main.cu:8,115);
//   * F_s is stored at the 0-based sample index (fself.cu:20, main.cu:772);
//   * the MSD is averaged over configs like the other observables
//     (main.cu:796-797 overwrote it);
//   * an instant quench (t0 == tf) gives Tf instead of NaN (set_temp.cu:8-11).
// The potential is the reference's active wca_modified (cutoff sigma,
// potentials.cu:66-76) unless BD_POTENTIAL selects standard WCA (cutoff
// 2^(1/6) sigma) or one of the 49-50 potentials (potentials.cu:9-52).
#include <sys/stat.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bd_abi.h"

namespace {

const double kPiRef = 3.1415927410125732;  // (double)3.14159266f, encabezados.h:25

void die(const std::string& msg, int code = 1) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::fflush(stdout);
  std::exit(code);
}

void check(int rc, const char* what) {
  if (rc != BD_OK) die(std::string(what) + ": " + bd_strerror(rc) + ": " + bd_last_error());
}

bool read_line(FILE* f, char* buf, int n) { return std::fgets(buf, n, f) != nullptr; }

// main.cu:14-53 (float arithmetic, as the reference)
std::vector<int> sampling_indexes(float minimal_decade, float run_time, float dt, int per) {
  const int last_power = (int)std::round(std::log10(run_time));
  const int initial_power = (int)std::round(std::log10(minimal_decade));
  const int decades = last_power - initial_power;
  std::vector<float> values(per + 1);
  const float r = std::pow(10.0f, 1.0f / (float)per);
  float x = 1.0f;
  for (int i = 0; i <= per; ++i) {
    values[i] = x;
    x *= r;
  }
  std::vector<int> idx(decades * per + 1);
  int c = 0;
  idx[0] = 0;
  for (int ipow = initial_power; ipow < last_power; ++ipow) {
    const float power = (float)std::pow(10.0, (double)ipow);
    for (int j = 0; j < per; ++j) idx[++c] = (int)std::round(values[j] * power / dt);
  }
  return idx;
}

// Lengths are read in FP64 (the reference reads them into float).
struct InitConfig {
  float phi_big = 0, phi_sml = 0, mass_big = 0, mass_sml = 0;
  double side = 0, sigma_big = 0, sigma_sml = 0;
  int n_big = 0, n_sml = 0, n_tot = 0;
  std::vector<double> xyz;
};

// main.cu:231-241,377-394
InitConfig read_init_config(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) {
    std::printf("Verify file path: %s", path.c_str());
    std::fflush(stdout);
    std::exit(-2);
  }
  InitConfig c;
  char line[512];
  read_line(f, line, sizeof line);
  std::sscanf(line, "%f %f", &c.phi_big, &c.phi_sml);
  read_line(f, line, sizeof line);
  std::sscanf(line, "%lf", &c.side);
  read_line(f, line, sizeof line);
  std::sscanf(line, "%d %d %d", &c.n_big, &c.n_sml, &c.n_tot);
  read_line(f, line, sizeof line);
  std::sscanf(line, "%lf %lf", &c.sigma_big, &c.sigma_sml);
  read_line(f, line, sizeof line);
  std::sscanf(line, "%f %f", &c.mass_big, &c.mass_sml);
  read_line(f, line, sizeof line);
  c.xyz.resize((size_t)3 * c.n_tot);
  for (int m = 0; m < c.n_tot; m++) {
    if (!read_line(f, line, sizeof line)) die("error: init config too short: " + path);
    int nn;
    double x, y, z;
    std::sscanf(line, "%d %lf %lf %lf", &nn, &x, &y, &z);
    if (nn != m) die("error: mm " + std::to_string(m) + "  nn " + std::to_string(nn) + " no match");
    c.xyz[3 * m] = x;
    c.xyz[3 * m + 1] = y;
    c.xyz[3 * m + 2] = z;
  }
  std::fclose(f);
  return c;
}

// main.cu:815-830 (and the format read above)
void write_init_config(const std::string& path, const InitConfig& c, const double* xyz) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) die("cannot write " + path, -2);
  std::fprintf(f, "%f %f    phi b, s\n", c.phi_big, c.phi_sml);
  std::fprintf(f, "%f    side\n", c.side);
  std::fprintf(f, "%d %d %d    ngr b, s, tot\n", c.n_big, c.n_sml, c.n_tot);
  std::fprintf(f, "%f %f    sigma b, s\n", c.sigma_big, c.sigma_sml);
  std::fprintf(f, "%f %f    mass b, s\n", c.mass_big, c.mass_sml);
  std::fprintf(f, "\n");
  for (int m = 0; m < c.n_tot; m++)
    std::fprintf(f, "%d %f %f %f\n", m, xyz[3 * m], xyz[3 * m + 1], xyz[3 * m + 2]);
  std::fclose(f);
}

double env_double(const char* name, double dflt) {
  const char* v = std::getenv(name);
  return v && *v ? std::strtod(v, nullptr) : dflt;
}

// BD_POTENTIAL: wca_modified (default), wca, wca49_50 or wca49_50_ao.  The
// 49-50 constants come from BD_POT_AMP and BD_POT_INV_T (default: the
// reference's commented AA = 50 (50/49)^50 and invT = 0.678575,
// get_forces.cu:46-48), the depletion branch's from BD_SIGMA_POL and
// BD_PHI_POL.
void set_potential_from_env(bd_params& p) {
  const char* v = std::getenv("BD_POTENTIAL");
  p.potential = BD_POT_WCA_MODIFIED;
  if (!v || !*v || std::strcmp(v, "wca_modified") == 0) return;
  if (std::strcmp(v, "wca") == 0) {
    p.potential = BD_POT_WCA;
  } else if (std::strcmp(v, "wca49_50") == 0) {
    p.potential = BD_POT_WCA49_50;
  } else if (std::strcmp(v, "wca49_50_ao") == 0) {
    p.potential = BD_POT_WCA49_50_AO;
  } else {
    std::fprintf(stderr, "unknown BD_POTENTIAL '%s' (wca_modified, wca, wca49_50, wca49_50_ao)\n",
                 v);
    std::exit(2);
  }
  if (p.potential == BD_POT_WCA49_50 || p.potential == BD_POT_WCA49_50_AO) {
    p.pot_amp = env_double("BD_POT_AMP", 50.0 * std::pow(50.0 / 49.0, 50.0));
    p.pot_inv_t = env_double("BD_POT_INV_T", 0.678575);
    p.sigma_pol = env_double("BD_SIGMA_POL", 0.0);
    p.phi_pol = env_double("BD_PHI_POL", 0.0);
  }
}

// ---------------------------------------------------------------- configs
uint64_t spread10(uint64_t v) {
  v &= 0x3ff;
  v = (v | (v << 16)) & 0x030000ffULL;
  v = (v | (v << 8)) & 0x0300f00fULL;
  v = (v | (v << 4)) & 0x030c30c3ULL;
  v = (v | (v << 2)) & 0x09249249ULL;
  return v;
}

// Z-order within each species so that neighbouring indices are neighbours
// in space (the kernels' gathers then share cache lines).
void spatial_sort(std::vector<double>& xyz, int n_big, int n_tot, double side) {
  auto sort_range = [&](int lo, int hi) {
    std::vector<std::pair<uint64_t, int>> key;
    for (int i = lo; i < hi; i++) {
      uint64_t q[3];
      for (int a = 0; a < 3; a++) {
        long v = (long)(xyz[3 * i + a] / side * 1024.0);
        q[a] = (uint64_t)std::min(1023L, std::max(0L, v));
      }
      key.push_back({spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2), i});
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int>& a, const std::pair<uint64_t, int>& b) {
                       return a.first < b.first;
                     });
    std::vector<double> tmp((size_t)3 * (hi - lo));
    for (int k = 0; k < hi - lo; k++)
      for (int a = 0; a < 3; a++) tmp[3 * k + a] = xyz[3 * key[k].second + a];
    std::copy(tmp.begin(), tmp.end(), xyz.begin() + 3 * (size_t)lo);
  };
  sort_range(0, n_big);
  sort_range(n_big, n_tot);
}

int make_configs(int argc, char** argv) {
  if (argc < 10) {
    std::fprintf(stderr,
                 "usage: %s --make-configs NBIG NSML PHI SIGMA_BIG SIGMA_SML NCONFIGS SEED DIR "
                 "[T EQ_STEPS DT]\n",
                 argv[0]);
    return 2;
  }
  const int nb = std::atoi(argv[2]), ns = std::atoi(argv[3]);
  const double phi = std::atof(argv[4]), sb = std::atof(argv[5]), ss = std::atof(argv[6]);
  const int nconf = std::atoi(argv[7]);
  const uint64_t seed = std::strtoull(argv[8], nullptr, 10);
  const std::string dir = argv[9];
  const double T = argc > 10 ? std::atof(argv[10]) : 1.0;
  const int eq_steps = argc > 11 ? std::atoi(argv[11]) : 0;
  const double dt = argc > 12 ? std::atof(argv[12]) : 1e-4;
  const int n = nb + ns;
  const double vol = kPiRef / 6.0 * (nb * sb * sb * sb + ns * ss * ss * ss);
  const double side = std::cbrt(vol / phi);
  // overlap removal: all configs at once, one replica each
  std::vector<bd_params> hp(nconf);
  for (int c = 0; c < nconf; c++) {
    bd_params& p = hp[c];
    std::memset(&p, 0, sizeof p);
    p.n_big = nb;
    p.n_sml = ns;
    p.potential = BD_POT_HARMONIC;
    p.ntags_big = p.ntags_sml = 24;
    p.side = side;
    p.sigma_big = sb;
    p.sigma_sml = ss;
    p.dt = 0.1;
    p.kappa = 1.0;
    p.seed = bd_lane_seed(seed, 1000 + c);
  }
  bd_sys* s = nullptr;
  check(bd_create(hp.data(), nconf, &s), "bd_create");
  std::vector<double> x((size_t)3 * n), f((size_t)3 * n), vir(n), pot(n);
  for (int c = 0; c < nconf; c++) {
    check(bd_random_positions(n, side, bd_lane_seed(seed, c), x.data()), "bd_random_positions");
    check(bd_set_positions(s, c, x.data()), "bd_set_positions");
  }
  const double target = 0.25 * 1.0 * std::pow(1e-3 * std::min(sb, ss), 2.0);
  int done = 0;
  for (;;) {
    check(bd_step(s, 256), "bd_step");
    done += 256;
    double worst = 0.0;
    for (int c = 0; c < nconf; c++) {
      check(bd_get_forces(s, c, f.data(), vir.data(), pot.data()), "bd_get_forces");
      worst = std::max(worst, *std::max_element(pot.begin(), pot.end()));
    }
    if (worst < target) break;
    if (done > 400000) die("overlaps remain after overlap removal");
  }
  std::vector<std::vector<double>> conf(nconf, std::vector<double>((size_t)3 * n));
  for (int c = 0; c < nconf; c++) check(bd_get_positions(s, c, conf[c].data()), "get");
  bd_destroy(s);
  std::printf("overlap removal: %d steps\n", done);
  if (eq_steps > 0) {
    std::vector<bd_params> ep(nconf);
    for (int c = 0; c < nconf; c++) {
      ep[c] = hp[c];
      set_potential_from_env(ep[c]);
      ep[c].ntags_big = ep[c].ntags_sml = 12;
      ep[c].dt = dt;
      ep[c].kappa = 0.0;
      ep[c].T0 = ep[c].Tf = T;
      ep[c].seed = bd_lane_seed(seed, 2000 + c);
    }
    check(bd_create(ep.data(), nconf, &s), "bd_create");
    for (int c = 0; c < nconf; c++) check(bd_set_positions(s, c, conf[c].data()), "set");
    check(bd_step(s, eq_steps), "bd_step");
    for (int c = 0; c < nconf; c++) check(bd_get_positions(s, c, conf[c].data()), "get");
    bd_destroy(s);
    std::printf("equilibration: %d steps at T=%g\n", eq_steps, T);
  }
  for (size_t k = 1; k <= dir.size(); k++)  // every missing level, as mkdir -p
    if (k == dir.size() || dir[k] == '/') mkdir(dir.substr(0, k).c_str(), 0755);
  InitConfig ic;
  ic.n_big = nb;
  ic.n_sml = ns;
  ic.n_tot = n;
  ic.side = side;
  ic.sigma_big = sb;
  ic.sigma_sml = ss;
  ic.mass_big = ic.mass_sml = 1.0f;
  ic.phi_big = (float)(kPiRef / 6.0 * nb * sb * sb * sb / (side * side * side));
  ic.phi_sml = (float)(kPiRef / 6.0 * ns * ss * ss * ss / (side * side * side));
  for (int c = 0; c < nconf; c++) {
    spatial_sort(conf[c], nb, n, side);
    write_init_config(dir + "/init_config_" + std::to_string(c + 1), ic, conf[c].data());
  }
  std::printf("wrote %d configs to %s (N=%d, L=%f)\n", nconf, dir.c_str(), n, side);
  return 0;
}

}  // namespace

// g(r) normalisation (main.cu:849-877) and gder_k.out files (main.cu:885-900).
void write_gders(const std::string& dir, std::vector<double>& gders, int nsamples, int nbins,
                 double range, const InitConfig& c0, int n_configs) {
  const double bin = range / (double)nbins;
  const double volume = (double)c0.side * c0.side * c0.side;
  const double xb = c0.n_big, xs = c0.n_sml;
  const double sbs = 0.5 * ((double)c0.sigma_big + (double)c0.sigma_sml);
  auto vfree = [&](double sg) { return volume - (4.0 / 3.0) * kPiRef * sg * sg * sg; };
  for (int k = 0; k < nsamples; k++) {
    double* gk = &gders[(size_t)k * 4 * nbins];
    for (int nb = 0; nb < nbins; nb++) {
      const double xnb = nb;
      const double shell = 4.0 * kPiRef * bin * bin * bin * ((1.0 / 3.0) + xnb * xnb + xnb);
      gk[nb] = gk[nb] / shell / (xb * (xb - 1.0) / vfree(c0.sigma_big)) / n_configs;
      gk[nbins + nb] = gk[nbins + nb] / shell / (xb * xs / vfree(sbs)) / n_configs;
      gk[2 * nbins + nb] = gk[2 * nbins + nb] / shell / (xs * xb / vfree(sbs)) / n_configs;
      gk[3 * nbins + nb] = gk[3 * nbins + nb] / shell / (xs * (xs - 1.0) / vfree(c0.sigma_sml)) /
                           n_configs;
    }
    FILE* f = std::fopen((dir + "gder_" + std::to_string(k) + ".out").c_str(), "w");
    if (!f) die("Verify file path: " + dir, -2);
    for (int nb = 0; nb < nbins; nb++)
      std::fprintf(f, "%f  %f  %f  %f  %f\n", (0.5 + nb) * bin, gk[nb], gk[nbins + nb],
                   gk[2 * nbins + nb], gk[3 * nbins + nb]);
    std::fclose(f);
  }
}

// The effective-force workflow of run-system.sh:118-145 (its bd_aging.bin is
// not part of the reference): the same protocol as the aging run with the
// configs directory line and a force_nbins / force_range line, the aging run
// of main.cu, and the Contracting_Vector3D force contraction
// (Contracting_Vector3D.cu:11-304) accumulated over every sample of every
// config on the GPU.  Writes results/gder_k.out, energies_averaged.out,
// pressure.out and force_contraction.out (bin centre, contracted force).
int run_system() {
  FILE* fp = std::fopen("bd_aging.data", "r");
  if (!fp) fp = stdin;
  char line[256], type[40] = {0}, sampling[40] = {0}, cfgdir[200] = {0};
  float T0 = 0, Tf = 0, t0 = 0, tf = 0, period = 1, gr_range = 0;
  double dt = 0, trans_time = 0, run_time = 0, minimal_decade = 0, force_range = 0;
  int ngap_rescaling = 0, nsamples = 0, tagb = 0, tags = 0, n_configs = 0, gr_nbins = 0,
      force_nbins = 0, idum = 0;
  std::printf("temp_protocol ?\n");
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%39s", type);
  const bool sine = std::strcmp(type, "sine") == 0;
  read_line(fp, line, sizeof line);
  if (std::strcmp(type, "linear") == 0)
    std::sscanf(line, "%f  %f  %f  %f  %d", &T0, &Tf, &t0, &tf, &ngap_rescaling);
  else if (sine)
    std::sscanf(line, "%f  %f  %f  %f  %f  %d", &T0, &Tf, &t0, &tf, &period, &ngap_rescaling);
  else
    die("unknown temperature protocol: " + std::string(type));
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%lf", &dt);
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%39s", sampling);
  const bool log_mode = std::strcmp(sampling, "log") == 0;
  read_line(fp, line, sizeof line);
  if (log_mode)
    std::sscanf(line, "%lf  %lf %lf", &trans_time, &minimal_decade, &run_time);
  else
    std::sscanf(line, "%lf %lf %d", &trans_time, &run_time, &nsamples);
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%d %d", &tagb, &tags);
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%199s", cfgdir);  // "../configs/" (not used, as in the reference)
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%d", &n_configs);
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%d %f", &gr_nbins, &gr_range);
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%d %lf", &force_nbins, &force_range);
  read_line(fp, line, sizeof line);
  std::sscanf(line, "%d", &idum);
  if (fp != stdin) std::fclose(fp);
  if (n_configs < 1 || gr_nbins < 1 || force_nbins < 1 || !(force_range > 0.0))
    die("bad run parameters");
  const int niter = (int)(run_time / dt), ntrans = (int)(trans_time / dt);
  std::vector<int> sample_at;
  if (log_mode) {
    sample_at = sampling_indexes((float)minimal_decade, (float)run_time, (float)dt, 10);
  } else {
    if (nsamples < 1) die("nsamples must be >= 1");
    const int ngap = niter / nsamples;
    if (ngap < 1) die("error: ngap " + std::to_string(ngap));
    for (int ni = 0; ni < niter; ni += ngap) sample_at.push_back(ni);
  }
  nsamples = (int)sample_at.size();
  std::vector<InitConfig> cfg(n_configs);
  for (int i = 0; i < n_configs; i++) {
    std::string path = "../configs/init_config_" + std::to_string(i + 1);
    struct stat st;
    if (stat(path.c_str(), &st) != 0)
      path = "../configs/decade0/init_config_" + std::to_string(i + 1);
    cfg[i] = read_init_config(path);
  }
  std::vector<bd_params> params(n_configs);
  for (int i = 0; i < n_configs; i++) {
    bd_params& p = params[i];
    std::memset(&p, 0, sizeof p);
    p.n_big = cfg[i].n_big;
    p.n_sml = cfg[i].n_sml;
    set_potential_from_env(p);
    p.ntags_big = tagb;
    p.ntags_sml = tags;
    p.temp_protocol = sine ? BD_TEMP_SINE : BD_TEMP_LINEAR;
    p.side = cfg[i].side;
    p.sigma_big = cfg[i].sigma_big;
    p.sigma_sml = cfg[i].sigma_sml;
    p.dt = dt;
    p.T0 = T0;
    p.Tf = Tf;
    p.t0 = t0;
    p.tf = tf;
    p.period = period;
    p.seed = bd_lane_seed((uint64_t)(int64_t)idum, (uint64_t)i);
    p.first_step = -ntrans;
  }
  bd_sys* s = nullptr;
  check(bd_create(params.data(), n_configs, &s), "bd_create");
  for (int i = 0; i < n_configs; i++) check(bd_set_positions(s, i, cfg[i].xyz.data()), "set");
  bd_cv3d* cv = nullptr;
  check(bd_cv3d_create(force_nbins, force_range, &cv), "bd_cv3d_create");
  std::vector<double> gders((size_t)nsamples * 4 * gr_nbins, 0.0), e_b(nsamples, 0.0),
      e_s(nsamples, 0.0), e_T(nsamples, 0.0), press(nsamples, 0.0), t_s(nsamples, 0.0);
  std::vector<uint64_t> counts((size_t)4 * gr_nbins);
  if (ntrans > 0) check(bd_step(s, ntrans), "bd_step");
  check(bd_mark_origin(s), "bd_mark_origin");
  int64_t ni_next = 0;
  for (int k = 0; k < nsamples; k++) {
    const int target = sample_at[k];
    const int64_t need = (int64_t)target + 1 - ni_next;
    if (need > 0) check(bd_step(s, (int32_t)need), "bd_step");
    ni_next += need > 0 ? need : 0;
    for (int r = 0; r < n_configs; r++) {
      bd_obs o;
      check(bd_observables(s, r, 1.0, &o), "bd_observables");
      e_b[k] += o.epot_big;
      e_s[k] += o.epot_sml;
      e_T[k] = o.temperature;
      press[k] += o.z;
      std::fill(counts.begin(), counts.end(), 0);
      check(bd_gder(s, r, gr_nbins, gr_range, counts.data()), "bd_gder");
      for (size_t q = 0; q < counts.size(); q++)
        gders[(size_t)k * 4 * gr_nbins + q] += (double)counts[q];
      check(bd_cv3d_add(cv, s, r, 1), "bd_cv3d_add");
    }
    t_s[k] = dt * (1.0 + (double)target);
    std::printf("run ni %d/%d  --  print %d/%d -- temp %.3f\n", target, niter, k + 1, nsamples,
                e_T[k]);
    std::fflush(stdout);
  }
  bd_destroy(s);
  check(bd_cv3d_solve(cv), "bd_cv3d_solve");
  std::vector<double> force((size_t)force_nbins);
  check(bd_cv3d_get(cv, force.data()), "bd_cv3d_get");
  bd_cv3d_destroy(cv);
  std::printf("printing files\n");
  const std::string dir = "results/";
  write_gders(dir, gders, nsamples, gr_nbins, gr_range, cfg[0], n_configs);
  auto open_out = [&](const std::string& name) {
    FILE* f = std::fopen((dir + name).c_str(), "w");
    if (!f) die("Verify file path: " + dir + name, -2);
    return f;
  };
  FILE* f = open_out("energies_averaged.out");
  for (int k = 0; k < nsamples; k++)
    std::fprintf(f, "%f  %f  %f  %f\n", t_s[k], e_b[k] / n_configs, e_s[k] / n_configs, e_T[k]);
  std::fclose(f);
  f = open_out("pressure.out");
  for (int k = 0; k < nsamples; k++) std::fprintf(f, "%f  %f\n", t_s[k], press[k] / n_configs);
  std::fclose(f);
  f = open_out("force_contraction.out");
  const double fb = force_range / force_nbins;
  for (int b = 0; b < force_nbins; b++) std::fprintf(f, "%f  %.10e\n", (0.5 + b) * fb, force[b]);
  std::fclose(f);
  std::printf("TERMINADO\n");
  return 0;
}

int main(int argc, char** argv) {
  {
    bd_params probe;  // an unknown BD_POTENTIAL ends the run before any work
    std::memset(&probe, 0, sizeof probe);
    set_potential_from_env(probe);
  }
  if (argc > 1 && std::strcmp(argv[1], "--make-configs") == 0) return make_configs(argc, argv);
  if (argc > 1 && std::strcmp(argv[1], "--system") == 0) return run_system();
  // --aging: in log mode with decade dumps, also measure F_s and the MSD from
  // every dumped waiting time within this one run (results/decadeK/) instead
  // of a reheat-and-restart run per decade
  const bool aging = argc > 1 && std::strcmp(argv[1], "--aging") == 0;

  // ---------------------------------------------------------------- protocol
  FILE* fp_data = std::fopen("wca_aging_linear.data", "r");
  if (!fp_data) fp_data = stdin;
  char renglon[200];
  char temperature_protocol[40] = {0}, sample_type[40] = {0};
  float T0 = 0, Tf = 0, t0 = 0, tf = 0, period = 1, range_gder = 0, qmax = 0;
  int ngap_rescaling = 0, nsamples = 0, ntags_big = 0, ntags_sml = 0, n_configs = 0,
      nbins_gder = 0, idum = 0;
  int print_decades_configs = 0, number_of_tws = 0, current_decade = 0;
  double dt = 0, trans_time = 0, run_time = 0, minimal_decade = 0;
  std::printf("temp_protocol ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%39s", temperature_protocol);
  std::printf("temp_settings ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  const bool sine = std::strcmp(temperature_protocol, "sine") == 0;
  if (std::strcmp(temperature_protocol, "linear") == 0)
    std::sscanf(renglon, "%f  %f  %f  %f  %d", &T0, &Tf, &t0, &tf, &ngap_rescaling);
  else if (sine)
    std::sscanf(renglon, "%f  %f  %f  %f  %f  %d", &T0, &Tf, &t0, &tf, &period, &ngap_rescaling);
  else
    die("unknown temperature protocol: " + std::string(temperature_protocol));
  std::printf("dt ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%lf", &dt);
  std::printf("sample type (linear, log) ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%39s", sample_type);
  const bool log_mode = std::strcmp(sample_type, "log") == 0;
  if (log_mode) {
    std::printf("trans_time, minimal_decade, run_time, nsamples (taken in run)\n");
    read_line(fp_data, renglon, sizeof renglon);
    std::sscanf(renglon, "%lf  %lf %lf", &trans_time, &minimal_decade, &run_time);
    std::printf("print decade configs?\n");
    read_line(fp_data, renglon, sizeof renglon);
    std::sscanf(renglon, "%d  %d  %d", &print_decades_configs, &number_of_tws, &current_decade);
  } else {
    std::printf("trans_time, run_time, nsamples (taken in run)\n");
    read_line(fp_data, renglon, sizeof renglon);
    std::sscanf(renglon, "%lf %lf %d", &trans_time, &run_time, &nsamples);
  }
  std::printf("ntags: big, sml ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%d %d", &ntags_big, &ntags_sml);
  std::printf("n_configs ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%d", &n_configs);
  std::printf("nbins_gder, range_gder ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%d %f", &nbins_gder, &range_gder);
  std::printf("qmax ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%f", &qmax);
  std::printf("idum ?\n");
  read_line(fp_data, renglon, sizeof renglon);
  std::sscanf(renglon, "%d", &idum);
  if (fp_data != stdin) std::fclose(fp_data);
  std::printf("Starts simulation_3D_2SP Device\n");
  if (n_configs < 1) die("n_configs must be >= 1");

  // ---------------------------------------------------------------- sampling
  std::vector<int> sample_at;  // sample step indices ni (>= 0)
  const int niter = (int)(run_time / dt);
  const int ntrans = (int)(trans_time / dt);
  const int points_per_decade = 10;
  if (log_mode) {
    sample_at = sampling_indexes((float)minimal_decade, (float)run_time, (float)dt,
                                 points_per_decade);
    nsamples = (int)sample_at.size();
  } else {
    if (nsamples < 1) die("nsamples must be >= 1");
    const int ngap = niter / nsamples;
    if (ngap < 1) die("error: ngap " + std::to_string(ngap));
    for (int ni = 0; ni < niter; ni += ngap) sample_at.push_back(ni);
    nsamples = (int)sample_at.size();
  }

  // ---------------------------------------------------------------- configs
  std::vector<InitConfig> cfg(n_configs);
  for (int i = 0; i < n_configs; i++) {
    const std::string path =
        log_mode ? "../configs/decade" + std::to_string(current_decade) + "/init_config_" +
                       std::to_string(i + 1)
                 : "../configs/init_config_" + std::to_string(i + 1);
    cfg[i] = read_init_config(path);
  }
  const InitConfig& c0 = cfg[0];
  std::vector<bd_params> params(n_configs);
  for (int i = 0; i < n_configs; i++) {
    bd_params& p = params[i];
    std::memset(&p, 0, sizeof p);
    p.n_big = cfg[i].n_big;
    p.n_sml = cfg[i].n_sml;
    set_potential_from_env(p);
    p.ntags_big = ntags_big;
    p.ntags_sml = ntags_sml;
    p.temp_protocol = sine ? BD_TEMP_SINE : BD_TEMP_LINEAR;
    p.side = cfg[i].side;
    p.sigma_big = cfg[i].sigma_big;
    p.sigma_sml = cfg[i].sigma_sml;
    p.dt = dt;
    p.T0 = T0;
    p.Tf = Tf;
    p.t0 = t0;
    p.tf = tf;
    p.period = period;
    p.seed = bd_lane_seed((uint64_t)(int64_t)idum, (uint64_t)i);
    p.first_step = -ntrans;
  }
  bd_geometry g;
  check(bd_compute_geometry(&params[0], &g), "geometry");
  FILE* fp_bitac = std::fopen("bitacora", "w");
  if (fp_bitac) {
    std::fprintf(fp_bitac, "side  %f  ngr b s tot  %d  %d  %d\n", c0.side, c0.n_big, c0.n_sml,
                 c0.n_tot);
    std::fprintf(fp_bitac, "ncell b, s  %d  %d\n", g.ncell_big, g.ncell_sml);
    std::fprintf(fp_bitac, "cell_side b, s  %f  %f\n", g.cell_side_big, g.cell_side_sml);
    std::fprintf(fp_bitac, "ntags b, s  %d  %d\n", ntags_big, ntags_sml);
    std::fprintf(fp_bitac, "ranges bb, bs  %d  %d\n", g.nrange_bb, g.nrange_bs);
    std::fprintf(fp_bitac, "ranges sb, ss  %d  %d\n", g.nrange_sb, g.nrange_ss);
    std::fprintf(fp_bitac, "niter, ntrans, nsamples %d %d %d\n", niter, ntrans, nsamples);
    std::fprintf(fp_bitac, "replicas (configs run concurrently) %d\n", n_configs);
    std::fflush(fp_bitac);
  }
  bd_sys* s = nullptr;
  check(bd_create(params.data(), n_configs, &s), "bd_create");
  for (int i = 0; i < n_configs; i++) check(bd_set_positions(s, i, cfg[i].xyz.data()), "set");

  // ---------------------------------------------------------------- accumulators
  std::vector<double> gders((size_t)nsamples * 4 * nbins_gder, 0.0);
  std::vector<double> fself(nsamples, 0.0), e_b(nsamples, 0.0), e_s(nsamples, 0.0),
      e_T(nsamples, 0.0), msd_b(nsamples, 0.0), msd_s(nsamples, 0.0), press(nsamples, 0.0),
      t_s(nsamples, 0.0);
  std::vector<uint64_t> counts((size_t)4 * nbins_gder);
  int number_of_tw = 1;
  // one-run aging: waiting-time origins (slot j = decade j + 1)
  const int n_origins = aging && log_mode && print_decades_configs
                            ? std::min(number_of_tws, (int)BD_MAX_ORIGINS)
                            : 0;
  int origins_set = 0;
  std::vector<double> t_origin(n_origins, 0.0);
  std::vector<std::vector<double>> o_lag(n_origins), o_fs(n_origins), o_msdb(n_origins),
      o_msds(n_origins);
  std::vector<double> o_out;

  // transient, then ni == 0 resets the aging clock (main.cu:497-506)
  if (ntrans > 0) check(bd_step(s, ntrans), "bd_step");
  check(bd_mark_origin(s), "bd_mark_origin");
  int64_t ni_next = 0;  // index of the next update
  for (int k = 0; k < nsamples; k++) {
    const int target = sample_at[k];
    if (target < ni_next - 1) continue;
    // state after the update + forces of step `target`
    const int64_t need = (int64_t)target + 1 - ni_next;
    if (need > 0) check(bd_step(s, (int32_t)need), "bd_step");
    ni_next += need > 0 ? need : 0;
    const double time = dt * (1.0 + (double)target);
    int ocb = 0, ocs = 0;
    for (int r = 0; r < n_configs; r++) {
      bd_obs o;
      check(bd_observables(s, r, qmax, &o), "bd_observables");
      e_b[k] += o.epot_big;
      e_s[k] += o.epot_sml;
      e_T[k] = o.temperature;
      fself[k] += o.fself_big;
      msd_b[k] += o.msd_big;
      msd_s[k] += o.msd_sml;
      press[k] += o.z;
      std::fill(counts.begin(), counts.end(), 0);
      check(bd_gder(s, r, nbins_gder, range_gder, counts.data()), "bd_gder");
      for (size_t q = 0; q < counts.size(); q++) gders[(size_t)k * 4 * nbins_gder + q] += (double)counts[q];
      if (r == 0) check(bd_max_occupancy(s, r, &ocb, &ocs), "bd_max_occupancy");
      // decade dumps (main.cu:809-835)
      if (log_mode && k % points_per_decade == 1 && print_decades_configs &&
          number_of_tw <= number_of_tws) {
        std::vector<double> pos((size_t)3 * cfg[r].n_tot);
        check(bd_get_positions(s, r, pos.data()), "get");
        const std::string path = "../configs/decade" + std::to_string(number_of_tw) +
                                 "/init_config_" + std::to_string(r + 1);
        write_init_config(path, cfg[r], pos.data());
      }
    }
    // one-run aging: F_s / MSD against every origin set so far, then a new
    // origin where this sample dumps a decade configuration
    if (origins_set > 0) {
      o_out.assign((size_t)n_configs * origins_set * 3, 0.0);
      check(bd_fself_origins(s, qmax, origins_set, o_out.data()), "bd_fself_origins");
      for (int j = 0; j < origins_set; j++) {
        double fs = 0, mb = 0, ms = 0;
        for (int r = 0; r < n_configs; r++) {
          const double* o = &o_out[((size_t)r * origins_set + j) * 3];
          fs += o[0];
          mb += o[1];
          ms += o[2];
        }
        o_lag[j].push_back(time - t_origin[j]);
        o_fs[j].push_back(fs / n_configs);
        o_msdb[j].push_back(mb / n_configs);
        o_msds[j].push_back(ms / n_configs);
      }
    }
    if (log_mode && k % points_per_decade == 1 && print_decades_configs &&
        number_of_tw <= number_of_tws) {
      if (number_of_tw - 1 < n_origins) {
        check(bd_set_origin(s, number_of_tw - 1), "bd_set_origin");
        t_origin[number_of_tw - 1] = time;
        origins_set = number_of_tw;
      }
      number_of_tw++;
    }
    t_s[k] = time;
    std::printf("run ni %d/%d  %.2f%%  --  print %d/%d  --  ocup b s  %d %d -- temp %.3f\n", target,
                niter, 100.0 * (double)target / (double)niter, k + 1, nsamples, ocb, ocs, e_T[k]);
    std::fflush(stdout);
  }
  bd_destroy(s);
  for (int i = 0; i < n_configs; i++) std::printf("Finished step %d of %d\n", i + 1, n_configs);

  // ---------------------------------------------------------------- normalise and write
  std::printf("printing files\n");
  const std::string dir = log_mode ? "results/decade" + std::to_string(current_decade) + "/"
                                   : std::string("results/");
  auto open_out = [&](const std::string& name) {
    FILE* f = std::fopen((dir + name).c_str(), "w");
    if (!f) {
      std::printf("Verify file path: %s", (dir + name).c_str());
      std::fflush(stdout);
      std::exit(-2);
    }
    return f;
  };
  write_gders(dir, gders, nsamples, nbins_gder, range_gder, c0, n_configs);
  FILE* f = open_out("energies_averaged.out");
  for (int k = 0; k < nsamples; k++)
    std::fprintf(f, "%f  %f  %f  %f\n", t_s[k], e_b[k] / n_configs, e_s[k] / n_configs, e_T[k]);
  std::fclose(f);
  f = open_out("fself.out");
  for (int k = 0; k < nsamples; k++) std::fprintf(f, "%f  %f\n", t_s[k], fself[k] / n_configs);
  std::fclose(f);
  f = open_out("msd_averaged.out");
  for (int k = 0; k < nsamples; k++)
    std::fprintf(f, "%f  %f  %f\n", t_s[k], msd_b[k] / n_configs, msd_s[k] / n_configs);
  std::fclose(f);
  f = open_out("pressure.out");
  for (int k = 0; k < nsamples; k++) std::fprintf(f, "%f  %f\n", t_s[k], press[k] / n_configs);
  std::fclose(f);
  for (int j = 0; j < n_origins; j++) {
    const std::string od = "results/decade" + std::to_string(j + 1) + "/";
    FILE* fo = std::fopen((od + "fself.out").c_str(), "w");
    FILE* fm = std::fopen((od + "msd_averaged.out").c_str(), "w");
    if (!fo || !fm) {
      std::printf("Verify file path: %s", od.c_str());
      std::fflush(stdout);
      std::exit(-2);
    }
    for (size_t k = 0; k < o_lag[j].size(); k++) {
      std::fprintf(fo, "%f  %f\n", o_lag[j][k], o_fs[j][k]);
      std::fprintf(fm, "%f  %f  %f\n", o_lag[j][k], o_msdb[j][k], o_msds[j][k]);
    }
    std::fclose(fo);
    std::fclose(fm);
  }
  if (fp_bitac) std::fclose(fp_bitac);
  std::printf("TERMINADO\n");
  return 0;
}

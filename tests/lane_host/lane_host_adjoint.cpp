// lane_host_adjoint.cpp -- runs the host build of a generated adjoint kernel
// (PFLANE_ADJOINT) over SoA state files, one cell at a time (test
// infrastructure, see lane_host_shim.h).
// usage: lane_host_adjoint <dir> <ncell> <ld> <dt> <jacobian 0|1> <npar> <nfield>
// <dir>/<field>.bin holds each input (FP64); wv.bin / ws.bin the weights;
// writes grad.out ([npar][ncell]), fcell.out, status.out.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static std::vector<char> slurp(const std::string& p) {
  std::vector<char> b;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  std::fseek(f, 0, SEEK_END);
  b.resize(std::ftell(f));
  std::fseek(f, 0, SEEK_SET);
  if (!b.empty() && std::fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
  std::fclose(f);
  return b;
}
static void spill(const std::string& p, const void* d, size_t n) {
  FILE* f = std::fopen(p.c_str(), "wb");
  std::fwrite(d, 1, n, f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 8) return 2;
  const std::string dir = argv[1];
  pflane::SensArgs a{};
  a.ncell = std::atoll(argv[2]);
  a.ld = std::atoll(argv[3]);
  a.dt = std::atof(argv[4]);
  a.jacobian = std::atoi(argv[5]);
  const int npar = std::atoi(argv[6]);
  const int nf = std::atoi(argv[7]);
  const char* names[] = {"pri_molal", "total", "pri_act_coef", "sec_act_coef", "ln_act_h2o",
                         "mnrl_volfrac", "mnrl_area", "porosity", "saturation", "den_kg",
                         "volume", "temperature", "wv", "ws"};
  const double** slots[] = {&a.pri_molal, &a.total, &a.pri_act_coef, &a.sec_act_coef,
                            &a.ln_act_h2o, &a.mnrl_volfrac, &a.mnrl_area, &a.porosity,
                            &a.saturation, &a.den_kg, &a.volume, &a.temperature, &a.wv, &a.ws};
  std::vector<std::vector<char>> buf(14);
  for (int k = 0; k < 14; ++k) {
    buf[k] = slurp(dir + "/" + names[k] + ".bin");
    *slots[k] = buf[k].empty() ? nullptr : (const double*)buf[k].data();
  }
  std::vector<double> grad((size_t)(npar > 0 ? npar : 1) * a.ncell), fcell(a.ncell);
  std::vector<int> status(a.ncell, -1);
  std::vector<double> lg((size_t)pflane::kSensSub * nf * a.ncell);
  a.grad = grad.data();
  a.ldo = a.ncell;
  a.fcell = fcell.data();
  a.status = status.data();
  a.log = lg.data();
  a.ldl = a.ncell;
  for (long long c = 0; c < a.ncell; ++c) {
    blockIdx.x = (unsigned)c;
    pfchem_lane_adjoint(a);
  }
  spill(dir + "/grad.out", grad.data(), grad.size() * sizeof(double));
  spill(dir + "/fcell.out", fcell.data(), fcell.size() * sizeof(double));
  spill(dir + "/status.out", status.data(), status.size() * sizeof(int));
  return 0;
}

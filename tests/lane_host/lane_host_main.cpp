// lane_host_main.cpp -- runs the host build of a generated one-lane kernel
// over SoA state files (test infrastructure, see lane_host_shim.h).
// usage: lane_host <dir> <ncell> <ld> <dt> <f32 0|1> <escalate 0|1>
// <dir>/<field>.bin holds each field (FP64, or int32 for the counters).
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
static void spill(const std::string& p, const std::vector<char>& b) {
  FILE* f = std::fopen(p.c_str(), "wb");
  std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 7) return 2;
  const std::string dir = argv[1];
  pflane::Args a{};
  a.ncell = std::atoll(argv[2]);
  a.ld = std::atoll(argv[3]);
  a.dt = std::atof(argv[4]);
  const int f32 = std::atoi(argv[5]);
  a.escalate = std::atoi(argv[6]);
  a.nxcd = 1;
  const char* names[] = {"pri_molal", "total", "immobile", "sec_molal", "pri_act_coef",
                         "sec_act_coef", "ln_act_h2o", "mnrl_volfrac", "mnrl_area", "mnrl_rate",
                         "porosity", "saturation", "den_kg", "volume", "temperature",
                         "num_iterations", "num_sub_steps", "status"};
  std::vector<std::vector<char>> buf(18);
  void** slots[] = {(void**)&a.pri_molal, (void**)&a.total, (void**)&a.immobile,
                    (void**)&a.sec_molal, (void**)&a.pri_act_coef, (void**)&a.sec_act_coef,
                    (void**)&a.ln_act_h2o, (void**)&a.mnrl_volfrac, (void**)&a.mnrl_area,
                    (void**)&a.mnrl_rate, (void**)&a.porosity, (void**)&a.saturation,
                    (void**)&a.den_kg, (void**)&a.volume, (void**)&a.temperature,
                    (void**)&a.num_iterations, (void**)&a.num_sub_steps, (void**)&a.status};
  for (int k = 0; k < 18; ++k) {
    buf[k] = slurp(dir + "/" + names[k] + ".bin");
    *slots[k] = buf[k].empty() ? nullptr : (void*)buf[k].data();
  }
  unsigned long long cursors[16 * 64] = {0};
  a.cursors = cursors;
  a.rescue = 0;
  if (f32) {
    pfchem_lane_step_f32(a);
    if (a.escalate && !buf[17].empty() && !std::getenv("PFCHEM_HOST_NO_RESCUE")) {
      // the FP64 pass over the list of handed-over cells, as the library
      // runs it after the FP32 pass
      std::vector<int> list;
      const int* st = (const int*)buf[17].data();
      for (long long c = 0; c < a.ncell; ++c)
        if (st[c] == 16) list.push_back((int)c);
      const int count = (int)list.size();
      list.push_back(0);  // a valid address when the list is empty
      cursors[0] = 0;
      a.rescue = 1;
      a.list = list.data();
      a.list_count = &count;
      a.list_min = 0;
      pfchem_lane_step_f64(a);
    }
  } else {
    pfchem_lane_step_f64(a);
  }
  for (int k = 0; k < 18; ++k)
    if (!buf[k].empty()) spill(dir + "/" + names[k] + ".out", buf[k]);
  return 0;
}

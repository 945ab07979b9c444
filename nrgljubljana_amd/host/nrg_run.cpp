// nrg_run.cpp - libnrgrun.so: the C ABI of a whole NRG calculation (include/nrgrun.h).
#include <dlfcn.h>

#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

#include "../../include/nrgrun.h"
#include "nrg/driver.hpp"

namespace {
thread_local std::string last_error;

std::string library_dir() {
  Dl_info info{};
  if (dladdr((void *)&nrg_run_dir, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    const auto s = p.find_last_of('/');
    return s == std::string::npos ? "." : p.substr(0, s);
  }
  return ".";
}

void write_timing(const std::string &dir, const NRG::Timing &t, nrg_ctx *ctx) {
  std::ofstream F(dir + "/nrg_timing.json");
  F << std::setprecision(9) << "{\"sections\": {";
  bool first = true;
  for (const auto &[k, v] : t.s) {
    F << (first ? "" : ", ") << "\"" << k << "\": " << v;
    first = false;
  }
  F << "}, \"nrg_steps\": [";
  for (size_t i = 0; i < t.nrg_steps.size(); i++) F << (i ? ", " : "") << t.nrg_steps[i].second;
  F << "], \"dm_steps\": [";
  for (size_t i = 0; i < t.dm_steps.size(); i++) F << (i ? ", " : "") << t.dm_steps[i].second;
  F << "], \"blocks\": [";
  for (size_t i = 0; i < t.blocks.size(); i++) {
    F << (i ? ", " : "") << "[";
    for (size_t k = 0; k < t.blocks[i].size(); k++) F << (k ? ", " : "") << t.blocks[i][k];
    F << "]";
  }
  F << "]";
  // the eigensolver autotuner's choices over the run's block histograms
  std::vector<char> buf(1 << 16);
  if (ctx && nrg_syev_tuning(ctx, buf.data(), (int)buf.size()) == NRG_OK) F << ", \"syev_tuning\": " << buf.data();
  F << "}\n";
}
}  // namespace

extern "C" int nrg_run_dir(const char *dir, const char *tables_dir, int device, const char *log_path) {
  try {
    if (!dir) throw std::invalid_argument("nrg_run_dir: dir is NULL");
    const std::string tables = tables_dir ? tables_dir : library_dir() + "/tables";
    std::ofstream logf;
    if (log_path) {
      logf.open(log_path);
      if (!logf) throw std::runtime_error(std::string("cannot open log file ") + log_path);
    }
    std::ostream &log = log_path ? (std::ostream &)logf : std::cout;
    {
      NRG::Device dev(device);
      NRG::NRGCalculation calc(dir, tables, dev, log);
      calc.run();
      write_timing(dir, calc.timing(), dev.ctx());
    }
    NRG::DevicePool::get().trim();   // every device buffer of the run has been released
    return 0;
  } catch (const std::exception &e) {
    last_error = e.what();
    NRG::DevicePool::get().trim();
    return 1;
  } catch (...) {
    last_error = "unknown exception";
    NRG::DevicePool::get().trim();
    return 1;
  }
}

extern "C" const char *nrg_run_error(void) { return last_error.c_str(); }

// main_t2c.cpp -- stand-alone driver with the command line of gput2cEXASIM:
//
//     gput2cEXASIM pdeapp.txt [model]
//
// It reads pdeapp.txt for the data paths the way the reference's t2c mode does
// (backend/Main/exasim.hpp:281-298; backend/Preprocessing/readpdeapp.cpp:187-260
// for the `key = value;` statements and :798-820 for datapath / modelnumber ->
// datain, dataout), solves the steady HDG problem in datain/ through the C ABI of
// libexahip.so, and writes dataout/out{udg,uhat}_np0.bin.  The model is named on
// the command line or by a `hipmodel = "..."` statement in pdeapp.txt (the
// reference links the generated model library instead).  Statements for the
// preprocessing stage (mesh, boundary expressions, ...) are read and ignored.
#include "exasim_hdg.h"

#include <cstdio>
#include <filesystem>
#include <fstream>
#include <map>
#include <regex>
#include <sstream>
#include <string>

namespace {

// `key = "text";` and `key = number;` statements; a statement may span lines and
// ends at ';' (or '}' for a list).  Lists and expressions are not needed here.
std::map<std::string, std::string> parse_pdeapp(const std::string& file, bool& ok) {
  std::map<std::string, std::string> kv;
  std::ifstream in(file);
  ok = static_cast<bool>(in);
  if (!ok) return kv;
  const std::regex str_re("([a-zA-Z0-9_]+)\\s*=\\s*\"([^\"]*)\"");
  const std::regex num_re("([a-zA-Z0-9_]+)\\s*=\\s*([^;\\[\\{]+)");
  std::string line, buf;
  while (std::getline(in, line)) {
    auto c = line.find('#');
    if (c != std::string::npos) line = line.substr(0, c);
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    buf += line;
    if (line.find(';') == std::string::npos && line.find('}') == std::string::npos) continue;
    std::smatch m;
    if (std::regex_search(buf, m, str_re)) kv[m[1]] = m[2];
    else if (std::regex_search(buf, m, num_re)) kv[m[1]] = m[2];
    buf.clear();
  }
  return kv;
}

std::string trim(const std::string& s) {
  auto a = s.find_first_not_of(" \t\r\n");
  auto b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: %s <pdeapp.txt> [model]\n", argv[0]);
    return 1;
  }
  const std::string app = argv[1];
  bool ok = false;
  auto kv = parse_pdeapp(app, ok);
  if (!ok) {
    std::fprintf(stderr, "Error: Input file %s does not exist.\n", app.c_str());
    return 1;
  }
  std::string model = argc > 2 ? argv[2] : (kv.count("hipmodel") ? kv["hipmodel"] : "");
  if (model.empty()) {
    std::fprintf(stderr, "Error: name the model on the command line or with hipmodel = \"...\"; in %s\n",
                 app.c_str());
    return 1;
  }
  namespace fs = std::filesystem;
  std::string datapath = kv.count("datapath") ? kv["datapath"] : "";
  if (datapath.empty()) {
    fs::path dir = fs::path(app).parent_path();
    datapath = dir.empty() ? fs::current_path().string() : dir.string();
  }
  int modelnumber = kv.count("modelnumber") ? std::stoi(trim(kv["modelnumber"])) : 0;
  std::string sfx = modelnumber > 0 ? std::to_string(modelnumber) : "";
  std::string datain = (fs::path(datapath) / ("datain" + sfx)).string();
  fs::path dataout = fs::path(datapath) / ("dataout" + sfx);
  std::error_code ec;
  fs::create_directories(dataout, ec);
  std::string prefix = (dataout / "out").string();
  std::printf("datain = %s\ndataout = %s\nmodel = %s\n", datain.c_str(), dataout.string().c_str(),
              model.c_str());

  exa_handle* h = nullptr;
  int rc = exa_create(datain.c_str(), prefix.c_str(), model.c_str(), -1, &h);
  if (rc != EXA_OK) {
    std::fprintf(stderr, "Error: %s\n", exa_last_error(h));
    exa_destroy(h);
    return rc;
  }
  exa_set_option(h, "verbose", 1);
  int newton = 0, gmres = 0;
  double norm = 0.0;
  rc = exa_solve(h, &newton, &gmres, &norm);
  if (rc == EXA_OK) rc = exa_write_output(h);
  if (rc != EXA_OK) {
    std::fprintf(stderr, "Error: %s\n", exa_last_error(h));
  } else {
    std::printf("Newton iterations %d, GMRES iterations %d, residual norm %g\n", newton, gmres, norm);
  }
  exa_destroy(h);
  return rc;
}

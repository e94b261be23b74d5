// datain.cpp -- readers of the reference's datain/*.bin files and writer of the
// dataout files.  Formats: backend/Discretization/readbinaryfiles.cpp (readers)
// and text2code/text2code/{readpdeapp,makemaster,writebinaryfiles}.cpp (writers).
// Every file is a float64 stream; integers are stored as doubles.
#include "common.hpp"

#include <cmath>
#include <fstream>

namespace exa {

namespace {

std::vector<double> read_all(const std::string& path) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) throw Error(2, "Unable to open file " + path);
  std::streamsize bytes = in.tellg();
  in.seekg(0);
  if (bytes % 8 != 0) throw Error(2, "File size is not a multiple of 8 bytes: " + path);
  std::vector<double> a(static_cast<size_t>(bytes / 8));
  if (bytes > 0 && !in.read(reinterpret_cast<char*>(a.data()), bytes))
    throw Error(2, "Read failed: " + path);
  return a;
}

// Cursor over a float64 stream: [lsize, nsize[lsize], arrays...]
struct Stream {
  const std::vector<double>& a;
  size_t pos = 0;
  std::string path;
  std::vector<double> take(size_t n) {
    if (pos + n > a.size()) throw Error(2, "Truncated file " + path);
    std::vector<double> out(a.begin() + pos, a.begin() + pos + n);
    pos += n;
    return out;
  }
  std::vector<int> take_int(size_t n) {
    std::vector<double> d = take(n);
    std::vector<int> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = static_cast<int>(std::lround(d[i]));
    return out;
  }
};

std::vector<int> header(Stream& s) {
  std::vector<int> l = s.take_int(1);
  if (l[0] <= 0 || l[0] > 1000) throw Error(2, "Bad header in " + s.path);
  return s.take_int(static_cast<size_t>(l[0]));
}

}  // namespace

// readbinaryfiles.cpp:65-173 (readappstruct)
AppData read_app(const std::string& path) {
  std::vector<double> a = read_all(path);
  Stream s{a, 0, path};
  std::vector<int> ns = header(s);
  if (ns.size() < 9) throw Error(2, "app.bin header too short");
  AppData app;
  app.ndims = s.take(ns[0]);
  app.flag = s.take(ns[1]);
  app.problem = s.take(ns[2]);
  app.uinf = s.take(ns[3]);
  app.dt = s.take(ns[4]);
  app.factor = s.take(ns[5]);
  app.physicsparam = s.take(ns[6]);
  app.solversparam = s.take(ns[7]);
  app.tau = s.take(ns[8]);
  if (app.ndims.size() < 14 || app.flag.size() < 13 || app.problem.size() < 28 ||
      app.solversparam.size() < 2 || app.tau.empty())
    throw Error(2, "app.bin arrays too short");
  return app;
}

// readbinaryfiles.cpp:208-277 (readmasterstruct)
MasterData read_master(const std::string& path) {
  std::vector<double> a = read_all(path);
  Stream s{a, 0, path};
  std::vector<int> ns = header(s);
  if (ns.size() < 15) throw Error(2, "master.bin header too short");
  std::vector<int> nd = s.take_int(ns[0]);
  MasterData m;
  m.nd = nd[0];
  m.elemtype = nd[1];
  m.nodetype = nd[2];
  m.porder = nd[3];
  m.pgauss = nd[4];
  m.npe = nd[5];
  m.npf = nd[6];
  m.nge = nd[7];
  m.ngf = nd[8];
  m.shapegt = s.take(ns[1]);
  m.shapegw = s.take(ns[2]);
  m.shapfgt = s.take(ns[3]);
  m.shapfgw = s.take(ns[4]);
  s.take(ns[5]);
  s.take(ns[6]);
  s.take(ns[7]);
  s.take(ns[8]);
  m.xpe = s.take(ns[9]);
  s.take(ns[10]);
  m.gwe = s.take(ns[11]);
  m.xpf = s.take(ns[12]);
  s.take(ns[13]);
  m.gwf = s.take(ns[14]);
  size_t ne1 = static_cast<size_t>(m.npe) * m.nge * (m.nd + 1);
  size_t nf1 = static_cast<size_t>(m.npf) * m.ngf * m.nd;
  if (m.shapegt.size() != ne1 || m.shapegw.size() != ne1 || m.shapfgt.size() != nf1 ||
      m.shapfgw.size() != nf1)
    throw Error(2, "master.bin shape tables have unexpected sizes");
  return m;
}

// readbinaryfiles.cpp:321-411 (readmeshstruct) -- the raw arrays buildConn needs.
MeshData read_mesh(const std::string& path) {
  std::vector<double> a = read_all(path);
  Stream s{a, 0, path};
  std::vector<int> ns = header(s);
  if (ns.size() < 29) throw Error(2, "mesh.bin header too short");
  std::vector<int> nd = s.take_int(ns[0]);
  MeshData m;
  m.nd = nd[0];
  m.ne = nd[1];
  m.nfe = nd[4];
  for (int k = 1; k <= 8; ++k) s.take(ns[k]);
  m.elempart = s.take_int(ns[9]);
  s.take(ns[10]);
  for (int k = 11; k <= 20; ++k) s.take(ns[k]);
  if (ns[1] != 0 || ns[21] != 0 || ns[22] != 0)
    throw Error(5, "mesh.bin with precomputed connectivity is not supported; "
                   "give the raw element-to-vertex array (text2code default)");
  s.take(ns[21]);
  s.take(ns[22]);
  m.perm = s.take_int(ns[23]);
  m.bf = s.take_int(ns[24]);
  m.cartgridpart = s.take_int(ns[25]);
  m.t = s.take_int(ns[26]);
  m.bcs = s.take_int(ns[27]);
  if (m.ne <= 0 || m.t.size() % m.ne) throw Error(2, "mesh.bin element array size mismatch");
  m.nve = static_cast<int>(m.t.size() / m.ne);
  if (static_cast<int>(m.bf.size()) != m.nfe * m.ne) throw Error(2, "mesh.bin bf size mismatch");
  if (static_cast<int>(m.elempart.size()) != m.ne) {
    m.elempart.resize(m.ne);
    for (int e = 0; e < m.ne; ++e) m.elempart[e] = e;
  }
  return m;
}

// readbinaryfiles.cpp:499-622 (readsolstruct)
SolData read_sol(const std::string& path, int npe, int nc, int ncu, int ne) {
  std::vector<double> a = read_all(path);
  Stream s{a, 0, path};
  std::vector<int> ns = header(s);
  if (ns.size() < 7) throw Error(2, "sol.bin header too short");
  s.take(ns[0]);
  SolData sol;
  sol.xdg = s.take(ns[1]);
  if (ns[1] % (npe * ne)) throw Error(2, "Input files are incorrect (xdg size)");
  std::vector<double> u = s.take(ns[2]);
  long npene = static_cast<long>(npe) * ne;
  if (!u.empty()) {
    if (static_cast<long>(u.size()) == npene * nc) {
      sol.udg = u;
      sol.udg_ncomp = nc;
    } else if (static_cast<long>(u.size()) == npene * ncu) {
      sol.udg = u;
      sol.udg_ncomp = ncu;
    } else {
      throw Error(2, "Input files are incorrect (udg size)");
    }
  }
  sol.odg = s.take(ns[3]);
  if (!sol.odg.empty()) {
    if (sol.odg.size() % npene) throw Error(2, "Input files are incorrect (vdg size)");
    sol.nco = static_cast<int>(sol.odg.size() / npene);
  }
  sol.wdg = s.take(ns[4]);  // w (the w equation's field), [npe, ncw, ne]
  if (sol.wdg.size() % npene) throw Error(2, "Input files are incorrect (wdg size)");
  sol.uh = s.take(ns[5]);
  return sol;
}

// solution.h:45-59 (open_and_write) + writearray
void write_output(const std::string& path, int a0, int a1, int a2, const std::vector<double>& data) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw Error(2, "Failed to open file: " + path);
  double hdr[3] = {double(a0), double(a1), double(a2)};
  out.write(reinterpret_cast<const char*>(hdr), sizeof(hdr));
  out.write(reinterpret_cast<const char*>(data.data()), data.size() * sizeof(double));
  if (!out) throw Error(2, "Write failed: " + path);
}

}  // namespace exa

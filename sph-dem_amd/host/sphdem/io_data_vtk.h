// io_data_vtk.h -- Cio_data_vtk for a one-rank run.cpp: the restart and
// output files of the reference (io_data_vtk.cpp:20-115 readRestart,
// :231-411 writeRestart, :413-720 writeOutput), the globals file
// (ioGlobals.cpp: a header line, then one line per output step) and the
// domain file (ioDomain.cpp).  VTK is not needed: the .vtu pieces are VTK XML
// UnstructuredGrid files with inline binary arrays, base64 of zlib blocks of
// 32 KiB with UInt32 headers (the encoding vtkXMLUnstructuredGridWriter
// writes with vtkZLibDataCompressor), and a .pvtu index beside them.  The
// reader also takes uncompressed binary and ascii arrays.  Positions are
// written as Float64 (the reference's vtkPoints default to Float32).
//
// Link with -lz.  Errors: a message on cerr, then exit(-1), as the reference.
#pragma once

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "customConstants.h"
#include "globalVars.h"
#include "particle.h"
#include "vect.h"

namespace sphdem_host {

namespace vtkxml {

inline std::string base64(const unsigned char* p, size_t n) {
  static const char* A = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string s;
  s.reserve(4 * ((n + 2) / 3));
  for (size_t i = 0; i < n; i += 3) {
    const uint32_t b = (uint32_t)p[i] << 16 | (i + 1 < n ? (uint32_t)p[i + 1] << 8 : 0) |
                       (i + 2 < n ? (uint32_t)p[i + 2] : 0);
    s += A[(b >> 18) & 63];
    s += A[(b >> 12) & 63];
    s += i + 1 < n ? A[(b >> 6) & 63] : '=';
    s += i + 2 < n ? A[b & 63] : '=';
  }
  return s;
}
inline std::vector<unsigned char> unbase64(const std::string& s) {
  std::vector<unsigned char> out;
  uint32_t acc = 0;
  int bits = 0;
  for (char ch : s) {
    int v;
    if (ch >= 'A' && ch <= 'Z') v = ch - 'A';
    else if (ch >= 'a' && ch <= 'z') v = ch - 'a' + 26;
    else if (ch >= '0' && ch <= '9') v = ch - '0' + 52;
    else if (ch == '+') v = 62;
    else if (ch == '/') v = 63;
    else continue;  // padding and white space
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((unsigned char)((acc >> bits) & 0xff));
    }
  }
  return out;
}
inline size_t b64len(size_t nbytes) { return 4 * ((nbytes + 2) / 3); }

// one array, compressed: header (nblocks, block size, last block size, the
// compressed sizes) and the blocks, each base64-encoded on its own
inline std::string encode(const void* data, size_t nbytes) {
  const uint32_t B = 32768;
  const unsigned char* p = (const unsigned char*)data;
  const uint32_t nb = nbytes ? (uint32_t)((nbytes + B - 1) / B) : 1;
  std::vector<uint32_t> hdr = {nb, B, nbytes ? (uint32_t)(nbytes - (size_t)(nb - 1) * B) : 0};
  std::vector<unsigned char> all;
  for (uint32_t k = 0; k < nb; k++) {
    const size_t len = nbytes ? std::min<size_t>(B, nbytes - (size_t)k * B) : 0;
    uLongf clen = compressBound(len);
    std::vector<unsigned char> c(clen);
    if (compress(c.data(), &clen, p + (size_t)k * B, len) != Z_OK) {
      std::cerr << "Cio_data_vtk: zlib compression failed" << std::endl;
      std::exit(-1);
    }
    hdr.push_back((uint32_t)clen);
    all.insert(all.end(), c.begin(), c.begin() + clen);
  }
  return base64((const unsigned char*)hdr.data(), hdr.size() * 4) + base64(all.data(), all.size());
}
inline std::vector<unsigned char> decode(const std::string& text, bool compressed) {
  std::string t;
  for (char ch : text)
    if (!isspace((unsigned char)ch)) t += ch;
  if (!compressed) {
    std::vector<unsigned char> raw = unbase64(t);
    if (raw.size() < 4) return {};
    uint32_t n;
    std::memcpy(&n, raw.data(), 4);
    return std::vector<unsigned char>(raw.begin() + 4, raw.begin() + 4 + n);
  }
  std::vector<unsigned char> first = unbase64(t.substr(0, b64len(12)));
  uint32_t nb;
  std::memcpy(&nb, first.data(), 4);
  const size_t hlen = b64len(4 * (3 + (size_t)nb));
  std::vector<unsigned char> hb = unbase64(t.substr(0, hlen));
  std::vector<uint32_t> hdr(3 + nb);
  std::memcpy(hdr.data(), hb.data(), 4 * hdr.size());
  std::vector<unsigned char> data = unbase64(t.substr(hlen));
  std::vector<unsigned char> out;
  size_t pos = 0;
  for (uint32_t k = 0; k < nb; k++) {
    const uLongf want = (k + 1 == nb && hdr[2]) ? hdr[2] : hdr[1];
    std::vector<unsigned char> blk(want);
    uLongf got = want;
    if (uncompress(blk.data(), &got, data.data() + pos, hdr[3 + k]) != Z_OK) {
      std::cerr << "Cio_data_vtk: a zlib block does not decompress" << std::endl;
      std::exit(-1);
    }
    out.insert(out.end(), blk.begin(), blk.begin() + got);
    pos += hdr[3 + k];
  }
  return out;
}

// the value of attribute `name` in the tag text `tag`
inline std::string attr(const std::string& tag, const std::string& name) {
  const std::string key = " " + name + "=\"";
  const size_t a = tag.find(key);
  if (a == std::string::npos) return "";
  const size_t b = tag.find('"', a + key.size());
  return tag.substr(a + key.size(), b - a - key.size());
}

struct Array {
  std::string type;
  int ncomp = 1;
  std::vector<double> v;  // every value widened to double (ints are exact)
};

// the Points array and the PointData arrays (by name) of a .vtu file's first
// piece
inline bool readVtu(const std::string& path, size_t& n, Array& points,
                    std::map<std::string, Array>& pd) {
  std::ifstream f(path);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  const bool compressed = s.find("compressor=") != std::string::npos;
  const size_t piece = s.find("<Piece");
  if (piece == std::string::npos) return false;
  n = std::strtoull(attr(s.substr(piece, s.find('>', piece) - piece), "NumberOfPoints").c_str(),
                    nullptr, 10);
  const size_t ptsAt = s.find("<Points>", piece);
  size_t pos = piece;
  while (true) {
    const size_t a = s.find("<DataArray", pos);
    if (a == std::string::npos) break;
    const size_t te = s.find('>', a);
    const size_t c = s.find("</DataArray>", te);
    const std::string tag = s.substr(a, te - a);
    const std::string body = s.substr(te + 1, c - te - 1);
    pos = c;
    Array arr;
    arr.type = attr(tag, "type");
    const std::string nc = attr(tag, "NumberOfComponents");
    arr.ncomp = nc.empty() ? 1 : std::atoi(nc.c_str());
    const std::string fmt = attr(tag, "format");
    if (fmt == "ascii") {
      std::istringstream in(body);
      double x;
      while (in >> x) arr.v.push_back(x);
    } else {
      std::vector<unsigned char> raw = decode(body, compressed);
      auto take = [&](auto tag0) {
        using T = decltype(tag0);
        const size_t m = raw.size() / sizeof(T);
        arr.v.resize(m);
        for (size_t i = 0; i < m; i++) {
          T x;
          std::memcpy(&x, raw.data() + i * sizeof(T), sizeof(T));
          arr.v[i] = (double)x;
        }
      };
      if (arr.type == "Float64") take(double());
      else if (arr.type == "Float32") take(float());
      else if (arr.type == "Int32") take(int32_t());
      else if (arr.type == "Int64") take(int64_t());
      else if (arr.type == "UInt8") take(uint8_t());
      else continue;
    }
    if (ptsAt != std::string::npos && a > ptsAt && a < s.find("</Points>", ptsAt)) points = arr;
    else if (a < ptsAt) pd[attr(tag, "Name")] = arr;
  }
  return true;
}

// a writer of one piece: arrays are added in order, then write()
class Piece {
 public:
  explicit Piece(size_t n) : n(n) {}
  template <class T>
  void add(const std::string& name, const std::string& type, int ncomp, const std::vector<T>& v) {
    std::ostringstream o;
    o << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\"";
    if (ncomp > 1) o << " NumberOfComponents=\"" << ncomp << "\"";
    o << " format=\"binary\">\n          " << encode(v.data(), v.size() * sizeof(T))
      << "\n        </DataArray>\n";
    arrays += o.str();
    layout.push_back({name, type, ncomp});
  }
  void write(const std::string& path, const std::vector<double>& r, const std::string& active) {
    std::ofstream f(path);
    if (!f) {
      std::cerr << "Cio_data_vtk: cannot write " << path << std::endl;
      std::exit(-1);
    }
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
         "byte_order=\"LittleEndian\" header_type=\"UInt32\" "
         "compressor=\"vtkZLibDataCompressor\">\n  <UnstructuredGrid>\n    <Piece NumberOfPoints=\""
      << n << "\" NumberOfCells=\"" << n << "\">\n      <PointData" << active << ">\n"
      << arrays << "      </PointData>\n      <Points>\n        <DataArray type=\"Float64\" "
      << "NumberOfComponents=\"3\" format=\"binary\">\n          "
      << encode(r.data(), r.size() * 8) << "\n        </DataArray>\n      </Points>\n"
      << "      <Cells>\n";
    std::vector<int32_t> idx(n), off(n);
    std::vector<uint8_t> ty(n, 1);
    for (size_t i = 0; i < n; i++) {
      idx[i] = (int32_t)i;
      off[i] = (int32_t)i + 1;
    }
    f << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"binary\">\n          "
      << encode(idx.data(), 4 * n) << "\n        </DataArray>\n"
      << "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"binary\">\n          "
      << encode(off.data(), 4 * n) << "\n        </DataArray>\n"
      << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"binary\">\n          "
      << encode(ty.data(), n) << "\n        </DataArray>\n"
      << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
  }
  void writeIndex(const std::string& path, const std::string& piece,
                  const std::string& active) const {
    std::ofstream f(path);
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" "
         "byte_order=\"LittleEndian\" compressor=\"vtkZLibDataCompressor\">\n"
         "  <PUnstructuredGrid GhostLevel=\"0\">\n    <PPointData"
      << active << ">\n";
    for (const auto& l : layout) {
      f << "      <PDataArray type=\"" << l.type << "\" Name=\"" << l.name << "\"";
      if (l.ncomp > 1) f << " NumberOfComponents=\"" << l.ncomp << "\"";
      f << "/>\n";
    }
    f << "    </PPointData>\n    <PPoints>\n      <PDataArray type=\"Float64\" "
         "NumberOfComponents=\"3\"/>\n    </PPoints>\n    <Piece Source=\""
      << piece << "\"/>\n  </PUnstructuredGrid>\n</VTKFile>\n";
  }

 private:
  struct L {
    std::string name, type;
    int ncomp;
  };
  size_t n;
  std::string arrays;
  std::vector<L> layout;
};

}  // namespace vtkxml

class Cio_data_vtk {
 public:
  Cio_data_vtk(std::string _filename, CglobalVars* g) { setFilename(_filename, g); }
  // base name of every file: baseNameNNNNNNN.pvtu (output), baseNameRestart
  // NNNNNNN.pvtu (restart), baseNameGlobalsN.dat, baseNameDomainN.dat; the
  // globals file is created with its header line if it does not exist
  // (ioGlobals.cpp:4-25)
  void setFilename(std::string _filename, CglobalVars* g) {
    filename = _filename;
    globalsFile = filename + "Globals" + std::to_string(g->mpiRank) + ".dat";
    domainFile = filename + "Domain.dat";
    std::ifstream in(globalsFile);
    if (!in) {
      std::ofstream fo(globalsFile);
      if (!fo) {
        std::cerr << "Could not open Global Vars file: " << globalsFile << std::endl;
        std::exit(-1);
      }
      fo << kGlobalsHeader << '\n';
    }
  }
  std::string getFilename() { return filename; }

  // readRestart (io_data_vtk.cpp:20-115): r, v, vhat, vort, norm1-3, dens, h,
  // the energies, u, mass, colour, iam and tag of every point of the piece
  void readRestart(int timestep, particleContainer* outPs, CglobalVars* g) {
    const std::string f = filename + "Restart" + step(timestep) + "_" +
                          std::to_string(g->mpiRank) + ".vtu";
    std::cout << "Reading file " << f << std::endl;
    size_t n = 0;
    vtkxml::Array pts;
    std::map<std::string, vtkxml::Array> pd;
    if (!vtkxml::readVtu(f, n, pts, pd)) {
      std::cerr << "Cio_data_vtk: cannot read " << f << std::endl;
      std::exit(-1);
    }
    std::cout << "Found " << n << " points in the file..." << std::endl;
    outPs->resize(n);
    auto vec = [&](const char* name, size_t i, vect& out) {
      auto it = pd.find(name);
      if (it != pd.end() && it->second.v.size() >= 3 * (i + 1))
        out = vect(it->second.v[3 * i], it->second.v[3 * i + 1], it->second.v[3 * i + 2]);
    };
    auto scl = [&](const char* name, size_t i, double& out) {
      auto it = pd.find(name);
      if (it != pd.end() && it->second.v.size() > i) out = it->second.v[i];
    };
    for (size_t i = 0; i < n; i++) {
      Cparticle& p = (*outPs)[i];
      p.r = vect(pts.v[3 * i], pts.v[3 * i + 1], pts.v[3 * i + 2]);
      vec("v", i, p.v);
      p.vhat = p.v;
      vec("vhat", i, p.vhat);
      vec("vort", i, p.vort);
      vec("norm1", i, p.norm1);
      vec("norm2", i, p.norm2);
      vec("norm3", i, p.norm3);
      scl("dens", i, p.dens);
      scl("h", i, p.h);
      scl("eBForce", i, p.eBForce);
      scl("eViscF", i, p.eViscF);
      scl("eViscB", i, p.eViscB);
      scl("eFF", i, p.eFF);
      scl("u", i, p.u);
      scl("mass", i, p.mass);
      scl("colour", i, p.colour);
      double t = 0.0;
      scl("iam", i, t);
      p.iam = (int)t;
      t = 0.0;
      scl("tag", i, t);
      p.tag = (int)t;
    }
  }
  // writeRestart (io_data_vtk.cpp:231-411): writeRestart's arrays in its
  // order, the piece of rank 0 and its .pvtu index
  void writeRestart(int timestep, particleContainer& ps, CglobalVars* g) {
    (void)g;
    const size_t n = ps.size();
    vtkxml::Piece P(n);
    std::vector<double> s(n), v3(3 * n), r(3 * n);
    auto scl = [&](const char* name, double Cparticle::*m) {
      for (size_t i = 0; i < n; i++) s[i] = ps[i].*m;
      P.add(name, "Float64", 1, s);
    };
    auto vec = [&](const char* name, vect Cparticle::*m) {
      for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) v3[3 * i + k] = (ps[i].*m)[k];
      P.add(name, "Float64", 3, v3);
    };
    scl("dens", &Cparticle::dens);
    scl("h", &Cparticle::h);
    scl("eBForce", &Cparticle::eBForce);
    scl("eViscF", &Cparticle::eViscF);
    scl("eViscB", &Cparticle::eViscB);
    scl("eFF", &Cparticle::eFF);
    scl("u", &Cparticle::u);
    scl("mass", &Cparticle::mass);
    scl("colour", &Cparticle::colour);
    std::vector<int32_t> iv(n);
    for (size_t i = 0; i < n; i++) iv[i] = ps[i].iam;
    P.add("iam", "Int32", 1, iv);
    for (size_t i = 0; i < n; i++) iv[i] = ps[i].tag;
    P.add("tag", "Int32", 1, iv);
    vec("v", &Cparticle::v);
    vec("vhat", &Cparticle::vhat);
    vec("vort", &Cparticle::vort);
    vec("norm1", &Cparticle::norm1);
    vec("norm2", &Cparticle::norm2);
    vec("norm3", &Cparticle::norm3);
    for (size_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) r[3 * i + k] = ps[i].r[k];
    const std::string stem = filename + "Restart" + step(timestep);
    const std::string active = " Scalars=\"iam\" Vectors=\"v\" Normals=\"norm1\"";
    P.write(stem + "_0.vtu", r, active);
    P.writeIndex(stem + ".pvtu", base(stem) + "_0.vtu", active);
    std::cout << "\tWriting " << n << " particles to file: " << stem << ".pvtu" << std::endl;
  }
  // writeOutput (io_data_vtk.cpp:413-720): the output arrays (FP32 fields,
  // tag as Int64) in writeOutput's order for the LIQ_DEM build
  void writeOutput(int timestep, particleContainer& ps, CglobalVars* g) {
    (void)g;
    const size_t n = ps.size();
    vtkxml::Piece P(n);
    std::vector<float> s(n), v3(3 * n);
    std::vector<double> r(3 * n);
    auto scl = [&](const char* name, double Cparticle::*m) {
      for (size_t i = 0; i < n; i++) s[i] = (float)(ps[i].*m);
      P.add(name, "Float32", 1, s);
    };
    auto vec = [&](const char* name, vect Cparticle::*m) {
      for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) v3[3 * i + k] = (float)(ps[i].*m)[k];
      P.add(name, "Float32", 3, v3);
    };
    scl("dens", &Cparticle::dens);
    scl("press", &Cparticle::press);
    std::vector<int64_t> tags(n);
    std::vector<int32_t> iams(n);
    for (size_t i = 0; i < n; i++) {
      tags[i] = ps[i].tag;
      iams[i] = ps[i].iam;
    }
    P.add("tag", "Int64", 1, tags);
    P.add("iam", "Int32", 1, iams);
    vec("v", &Cparticle::v);
    vec("vhat", &Cparticle::vhat);
#ifdef LIQ_DEM
    scl("porosity", &Cparticle::porosity);
    std::vector<float> zero(n, 0.0f);
    P.add("dporositydt", "Float32", 1, zero);  // not carried under the supported flags
    vec("fdrag", &Cparticle::fdrag);
    scl("shepSum", &Cparticle::shepSum);
#endif
    vec("vort", &Cparticle::vort);
    scl("h", &Cparticle::h);
    vec("ff", &Cparticle::ff);
    vec("fb", &Cparticle::fb);
    vec("fv", &Cparticle::fv);
    vec("fp", &Cparticle::fp);
    for (size_t i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) r[3 * i + k] = ps[i].r[k];
    const std::string stem = filename + step(timestep);
    const std::string active = " Scalars=\"dens\" Vectors=\"v\"";
    P.write(stem + "_0.vtu", r, active);
    P.writeIndex(stem + ".pvtu", base(stem) + "_0.vtu", active);
  }
  // writeGlobals / readGlobals (ioGlobals.cpp:42-50, 111-140): one line per
  // output step in the header's columns (the wall-clock columns 0), then
  // custom[]
  void writeGlobals(int outstep, CglobalVars* g) {
    std::ofstream fo(globalsFile, std::ios::app);
    fo << outstep << ' ' << g->sphStep << ' ' << g->time << ' ' << g->linMom[0] << ' '
       << g->linMom[1] << ' ' << g->angMom << ' ' << g->eKE << ' ' << g->eBForce << ' '
       << g->eElast << ' ' << g->eElastExact << ' ' << g->eViscF << ' ' << g->eViscB << ' '
       << g->eFF << ' ' << g->eTotal << " 0 0 0 0 0 0 " << g->maxV << ' ' << g->aveDens << ' '
       << g->varDens << ' ' << g->aveDensFromMass << ' ' << g->n << ' ' << g->nSph << ' '
       << g->maxF << ' ' << g->maxFF << ' ' << g->rmsFF << ' ' << g->edElastdt << ' '
       << g->edViscFdt << ' ' << g->edFFdt;
    for (int i = 0; i < 9; i++) fo << ' ' << g->custom[i];
    fo << '\n';
  }
  void readGlobals(int outstep, CglobalVars* g) {
    std::ifstream in(globalsFile);
    std::string line;
    std::getline(in, line);  // header
    while (std::getline(in, line)) {
      std::istringstream ss(line);
      int step = -1;
      double tmp;
      ss >> step;
      if (step != outstep) continue;
      ss >> g->sphStep >> g->time >> g->linMom[0] >> g->linMom[1] >> g->angMom >> g->eKE >>
          g->eBForce >> g->eElast >> g->eElastExact >> g->eViscF >> g->eViscB >> g->eFF >>
          g->eTotal >> tmp >> tmp >> tmp >> tmp >> tmp >> tmp >> g->maxV >> g->aveDens >>
          g->varDens >> g->aveDensFromMass >> g->n >> g->nSph >> g->maxF >> g->maxFF >>
          g->rmsFF >> g->edElastdt >> g->edViscFdt >> g->edFFdt;
      for (int i = 0; i < 9; i++) ss >> g->custom[i];
      return;
    }
  }
  // writeDomain / readDomain (ioDomain.cpp): the rank and its procDomain
  void writeDomain(int outstep, CglobalVars* g) {
    (void)outstep;
    std::ofstream fo(domainFile, std::ios::trunc);
    fo << g->mpiRank << ' ';
    for (int j = 0; j < 6; j++) fo << g->procDomain[j] << ' ';
    fo << '\n';
  }
  void readDomain(int outstep, CglobalVars* g) {
    (void)outstep;
    std::ifstream in(domainFile);
    if (!in) return;
    in >> g->mpiRank;
    for (int i = 0; i < 6; i++) in >> g->procDomain[i];
  }

  static constexpr const char* kGlobalsHeader =
      "# OutStep SPHStep Time linMomX linMomY angMom KE eBForce eElast eElastExact eViscF eViscB "
      "eFF eTotal wtTotalStep wtTotalForceCalc wtTotalDddtCalc wtTotalReset wtTotalUpdateDomain "
      "wtTotalMPI maxV aveDens varDens aveDensFromMass n nSph maxF maxFF rmsFF edElastdt "
      "edViscFdt edFFdt";

 private:
  static std::string step(int t) {
    char b[32];
    std::snprintf(b, sizeof b, "%7.7d", t);
    return b;
  }
  static std::string base(const std::string& path) {
    const size_t k = path.find_last_of('/');
    return k == std::string::npos ? path : path.substr(k + 1);
  }
  std::string filename, globalsFile, domainFile;
};

}  // namespace sphdem_host

using sphdem_host::Cio_data_vtk;

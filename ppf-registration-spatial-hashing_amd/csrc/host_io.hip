// host_io.hip -- host-side file formats and validation of the alignment
// driver: point clouds with normals from PLY (pcl::io::loadPLYFile into
// PointNormal, src/alignment.cpp:207-244), 4x4 ground-truth poses from .xf
// text files (operator>> of include/impl/util.hpp:95-104, used at
// src/alignment.cpp:300-311) and the pose distance of the validation,
// ht_dist (src/cuda/linalg.cu:9-20).  No device code.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ppf_objects.h"

using namespace ppfi;

namespace {

enum class PlyType { I8, U8, I16, U16, I32, U32, F32, F64, BAD };

PlyType ply_type(const std::string &t) {
  if (t == "char" || t == "int8") return PlyType::I8;
  if (t == "uchar" || t == "uint8") return PlyType::U8;
  if (t == "short" || t == "int16") return PlyType::I16;
  if (t == "ushort" || t == "uint16") return PlyType::U16;
  if (t == "int" || t == "int32") return PlyType::I32;
  if (t == "uint" || t == "uint32") return PlyType::U32;
  if (t == "float" || t == "float32") return PlyType::F32;
  if (t == "double" || t == "float64") return PlyType::F64;
  return PlyType::BAD;
}

int type_size(PlyType t) {
  switch (t) {
    case PlyType::I8:
    case PlyType::U8: return 1;
    case PlyType::I16:
    case PlyType::U16: return 2;
    case PlyType::I32:
    case PlyType::U32:
    case PlyType::F32: return 4;
    case PlyType::F64: return 8;
    default: return 0;
  }
}

struct PlyProp {
  std::string name;
  PlyType type = PlyType::BAD;
  bool list = false;
  PlyType count_type = PlyType::BAD;
};

struct PlyElement {
  std::string name;
  int64_t count = 0;
  std::vector<PlyProp> props;
};

struct PlyHeader {
  int format = 0;  // 0 ascii, 1 binary little endian, 2 binary big endian
  std::vector<PlyElement> elements;
  std::streampos data_start;
};

bool host_little_endian() {
  const uint16_t v = 1;
  return *reinterpret_cast<const uint8_t *>(&v) == 1;
}

double read_binary(std::istream &in, PlyType t, bool swap, bool &ok) {
  unsigned char b[8];
  const int sz = type_size(t);
  in.read(reinterpret_cast<char *>(b), sz);
  if (!in) {
    ok = false;
    return 0.0;
  }
  if (swap)
    for (int i = 0; i < sz / 2; ++i) std::swap(b[i], b[sz - 1 - i]);
  switch (t) {
    case PlyType::I8: return (double)(int8_t)b[0];
    case PlyType::U8: return (double)b[0];
    case PlyType::I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
    case PlyType::U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
    case PlyType::I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
    case PlyType::U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
    case PlyType::F32: { float v; std::memcpy(&v, b, 4); return v; }
    case PlyType::F64: { double v; std::memcpy(&v, b, 8); return v; }
    default: ok = false; return 0.0;
  }
}

int parse_header(std::ifstream &in, PlyHeader &h, std::string &err) {
  std::string line;
  if (!std::getline(in, line) || line.substr(0, 3) != "ply") {
    err = "not a PLY file";
    return PPF_ERR_ARG;
  }
  bool have_format = false;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ss(line);
    std::string kw;
    ss >> kw;
    if (kw == "format") {
      std::string f;
      ss >> f;
      if (f == "ascii") h.format = 0;
      else if (f == "binary_little_endian") h.format = 1;
      else if (f == "binary_big_endian") h.format = 2;
      else {
        err = "unknown PLY format " + f;
        return PPF_ERR_ARG;
      }
      have_format = true;
    } else if (kw == "element") {
      PlyElement e;
      ss >> e.name >> e.count;
      if (!ss || e.count < 0) {
        err = "bad element line: " + line;
        return PPF_ERR_ARG;
      }
      h.elements.push_back(e);
    } else if (kw == "property") {
      if (h.elements.empty()) {
        err = "property before any element";
        return PPF_ERR_ARG;
      }
      PlyProp p;
      std::string t;
      ss >> t;
      if (t == "list") {
        std::string ct, it;
        ss >> ct >> it >> p.name;
        p.list = true;
        p.count_type = ply_type(ct);
        p.type = ply_type(it);
      } else {
        p.type = ply_type(t);
        ss >> p.name;
      }
      if (p.type == PlyType::BAD || (p.list && p.count_type == PlyType::BAD)) {
        err = "unknown property type in: " + line;
        return PPF_ERR_ARG;
      }
      h.elements.back().props.push_back(p);
    } else if (kw == "end_header") {
      if (!have_format) {
        err = "PLY header without format";
        return PPF_ERR_ARG;
      }
      h.data_start = in.tellg();
      return PPF_OK;
    }
    // comment / obj_info and unknown keywords are skipped
  }
  err = "PLY header without end_header";
  return PPF_ERR_ARG;
}

// Reads the vertex element: x y z and nx ny nz (or normal_x ...); other
// properties and elements are skipped.  Missing normals are 0 (the
// default-constructed pcl::PointNormal fields).
int read_ply(const char *path, std::vector<float> &pts, std::vector<float> &nrm, std::string &err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    err = std::string("cannot open ") + path;
    return PPF_ERR_ARG;
  }
  PlyHeader h;
  int rc = parse_header(in, h, err);
  if (rc) return rc;
  const bool swap = (h.format == 1) != host_little_endian() && h.format != 0;
  for (const PlyElement &e : h.elements) {
    const bool vertex = e.name == "vertex";
    int col[6] = {-1, -1, -1, -1, -1, -1};
    if (vertex) {
      const char *names[6][2] = {{"x", "x"}, {"y", "y"}, {"z", "z"},
                                 {"nx", "normal_x"}, {"ny", "normal_y"}, {"nz", "normal_z"}};
      for (size_t k = 0; k < e.props.size(); ++k)
        for (int c = 0; c < 6; ++c)
          if (!e.props[k].list && (e.props[k].name == names[c][0] || e.props[k].name == names[c][1]))
            col[c] = (int)k;
      if (col[0] < 0 || col[1] < 0 || col[2] < 0) {
        err = "PLY vertex element without x, y, z";
        return PPF_ERR_ARG;
      }
      pts.assign((size_t)e.count * 3, 0.0f);
      nrm.assign((size_t)e.count * 3, 0.0f);
    }
    std::vector<double> vals(e.props.size());
    for (int64_t i = 0; i < e.count; ++i) {
      for (size_t k = 0; k < e.props.size(); ++k) {
        const PlyProp &p = e.props[k];
        bool ok = true;
        if (h.format == 0) {
          if (p.list) {
            double cnt = 0;
            in >> cnt;
            for (int64_t q = 0; q < (int64_t)cnt; ++q) {
              double dummy;
              in >> dummy;
            }
          } else {
            in >> vals[k];
          }
          ok = (bool)in;
        } else if (p.list) {
          const int64_t cnt = (int64_t)read_binary(in, p.count_type, swap, ok);
          for (int64_t q = 0; ok && q < cnt; ++q) read_binary(in, p.type, swap, ok);
        } else {
          vals[k] = read_binary(in, p.type, swap, ok);
        }
        if (!ok) {
          err = "PLY data ends early";
          return PPF_ERR_ARG;
        }
      }
      if (vertex) {
        for (int c = 0; c < 3; ++c) pts[3 * i + c] = (float)vals[col[c]];
        for (int c = 0; c < 3; ++c)
          if (col[3 + c] >= 0) nrm[3 * i + c] = (float)vals[col[3 + c]];
      }
    }
    if (vertex) return PPF_OK;
  }
  err = "PLY file without a vertex element";
  return PPF_ERR_ARG;
}

// Quaternion (w, x, y, z) of the 3x3 block of a row-major 4x4, in double
// (the larger of the trace and diagonal cases, as Eigen's
// Quaternion(Matrix3) does).
void quat_of_rot(const double R[3][3], double q[4]) {
  const double t = R[0][0] + R[1][1] + R[2][2];
  if (t > 0) {
    double s = std::sqrt(t + 1.0);
    q[0] = 0.5 * s;
    s = 0.5 / s;
    q[1] = (R[2][1] - R[1][2]) * s;
    q[2] = (R[0][2] - R[2][0]) * s;
    q[3] = (R[1][0] - R[0][1]) * s;
  } else {
    int i = 0;
    if (R[1][1] > R[0][0]) i = 1;
    if (R[2][2] > R[i][i]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    double s = std::sqrt(R[i][i] - R[j][j] - R[k][k] + 1.0);
    q[1 + i] = 0.5 * s;
    s = 0.5 / s;
    q[0] = (R[k][j] - R[j][k]) * s;
    q[1 + j] = (R[j][i] + R[i][j]) * s;
    q[1 + k] = (R[k][i] + R[i][k]) * s;
  }
}

}  // namespace

extern "C" {

int ppf_ply_read(const char *path, float *points, float *normals, int64_t capacity, int64_t *n_out) {
  if (!path || !n_out) {
    set_error("ppf_ply_read: NULL argument");
    return PPF_ERR_ARG;
  }
  std::vector<float> p, n;
  std::string err;
  const int rc = read_ply(path, p, n, err);
  if (rc) {
    set_error("ppf_ply_read: " + err);
    return rc;
  }
  const int64_t cnt = (int64_t)(p.size() / 3);
  *n_out = cnt;
  if (points || normals) {
    if (capacity < cnt) {
      set_error("ppf_ply_read: buffer holds fewer points than the file");
      return PPF_ERR_ARG;
    }
    if (points) std::memcpy(points, p.data(), p.size() * 4);
    if (normals) std::memcpy(normals, n.data(), n.size() * 4);
  }
  return PPF_OK;
}

int ppf_ply_write(const char *path, const float *points, const float *normals, int64_t n, int32_t binary) {
  if (!path || (!points && n > 0) || n < 0) {
    set_error("ppf_ply_write: NULL argument or n < 0");
    return PPF_ERR_ARG;
  }
  std::ofstream out(path, std::ios::binary);
  if (!out) {
    set_error(std::string("ppf_ply_write: cannot open ") + path);
    return PPF_ERR_ARG;
  }
  out << "ply\nformat " << (binary ? (host_little_endian() ? "binary_little_endian" : "binary_big_endian")
                                   : "ascii")
      << " 1.0\nelement vertex " << n << "\nproperty float x\nproperty float y\nproperty float z\n";
  if (normals) out << "property float nx\nproperty float ny\nproperty float nz\n";
  out << "end_header\n";
  for (int64_t i = 0; i < n; ++i) {
    if (binary) {
      out.write(reinterpret_cast<const char *>(points + 3 * i), 12);
      if (normals) out.write(reinterpret_cast<const char *>(normals + 3 * i), 12);
    } else {
      char buf[256];
      int len = std::snprintf(buf, sizeof buf, "%.9g %.9g %.9g", points[3 * i], points[3 * i + 1],
                              points[3 * i + 2]);
      out.write(buf, len);
      if (normals) {
        len = std::snprintf(buf, sizeof buf, " %.9g %.9g %.9g", normals[3 * i], normals[3 * i + 1],
                            normals[3 * i + 2]);
        out.write(buf, len);
      }
      out.put('\n');
    }
  }
  if (!out) {
    set_error("ppf_ply_write: write failed");
    return PPF_ERR_ARG;
  }
  return PPF_OK;
}

int ppf_xf_read(const char *path, float pose16[16]) {
  if (!path || !pose16) {
    set_error("ppf_xf_read: NULL argument");
    return PPF_ERR_ARG;
  }
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("ppf_xf_read: cannot open ") + path);
    return PPF_ERR_ARG;
  }
  for (int k = 0; k < 16; ++k) {
    if (!(in >> pose16[k])) {
      set_error(std::string("ppf_xf_read: fewer than 16 numbers in ") + path);
      return PPF_ERR_ARG;
    }
  }
  return PPF_OK;
}

int ppf_ht_dist(const float a16[16], const float b16[16], float out2[2]) {
  if (!a16 || !b16 || !out2) {
    set_error("ppf_ht_dist: NULL argument");
    return PPF_ERR_ARG;
  }
  // translation: norm(a_t - b_t) in float, ((x*x + y*y) + z*z)
  const float dx = a16[3] - b16[3], dy = a16[7] - b16[7], dz = a16[11] - b16[11];
  out2[0] = std::sqrt(dx * dx + dy * dy + dz * dz);
  // rotation: |angle| of AngleAxis(a_rot^-1 * b_rot): each rotation block as a
  // unit quaternion, q = conj(qa) * qb, angle = 2 atan2(|q.vec|, |q.w|)
  double A[3][3], B[3][3], qa[4], qb[4];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      A[r][c] = a16[4 * r + c];
      B[r][c] = b16[4 * r + c];
    }
  quat_of_rot(A, qa);
  quat_of_rot(B, qb);
  const double na = std::sqrt(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]);
  const double nb = std::sqrt(qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2] + qb[3] * qb[3]);
  for (int k = 0; k < 4; ++k) {
    qa[k] /= na;
    qb[k] /= nb;
  }
  const double w = qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2] + qa[3] * qb[3];
  const double x = qa[0] * qb[1] - qa[1] * qb[0] - qa[2] * qb[3] + qa[3] * qb[2];
  const double y = qa[0] * qb[2] + qa[1] * qb[3] - qa[2] * qb[0] - qa[3] * qb[1];
  const double z = qa[0] * qb[3] - qa[1] * qb[2] + qa[2] * qb[1] - qa[3] * qb[0];
  out2[1] = (float)(2.0 * std::atan2(std::sqrt(x * x + y * y + z * z), std::fabs(w)));
  return PPF_OK;
}

}  // extern "C"

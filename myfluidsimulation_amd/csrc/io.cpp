// io.cpp -- host-side file formats: STL input and VTK legacy ASCII output.
//
// STL (stl_loader.cpp:24-122): a file whose first five bytes are "solid" is
// parsed as ASCII (one triangle per three "vertex x y z" lines); if that
// yields no triangle it is re-read as binary (80-byte header, uint32 count,
// 50-byte records whose first 12 bytes, the normal, are ignored).  A short
// binary record is a failure.  Numbers go through the same C++ stream
// extraction as the reference, so every vertex has the reference's bits.
//
// VTK (moment_vtk_writer.cu:9-49): STRUCTURED_POINTS header, one value (or
// one "u v w" triple) per line, each number printed as std::ostream prints a
// float by default: "%.6g".  Formatting runs on several host threads, each
// filling its own slice of the text; the slices are written in order.  The
// binary variant (an extension) writes the same header with BINARY and the
// values as big-endian float32, the legacy format's binary encoding: exact
// values, about a third of the size, no formatting.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace fs {

namespace {

bool stl_ascii(const std::string& text, std::vector<float>& xyz) {
  xyz.clear();
  std::istringstream in(text);
  std::string line;
  float corner[9];
  int k = 0;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string word;
    ls >> word;
    if (word != "vertex") continue;
    float x = 0, y = 0, z = 0;
    ls >> x >> y >> z;
    corner[3 * k + 0] = x;
    corner[3 * k + 1] = y;
    corner[3 * k + 2] = z;
    if (++k == 3) {
      xyz.insert(xyz.end(), corner, corner + 9);
      k = 0;
    }
  }
  return !xyz.empty();
}

bool stl_binary(const std::string& bytes, std::vector<float>& xyz) {
  xyz.clear();
  if (bytes.size() < 84) return false;
  uint32_t ntri = 0;
  std::memcpy(&ntri, bytes.data() + 80, 4);
  size_t pos = 84;
  xyz.reserve(size_t(std::min<uint64_t>(ntri, bytes.size() / 50)) * 9);
  for (uint32_t t = 0; t < ntri; ++t) {
    if (pos + 50 > bytes.size()) return false;
    float rec[12];
    std::memcpy(rec, bytes.data() + pos, sizeof(rec));
    xyz.insert(xyz.end(), rec + 3, rec + 12);
    pos += 50;
  }
  return true;
}

}  // namespace

bool read_stl(const char* path, std::vector<float>& xyz) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream all;
  all << f.rdbuf();
  const std::string bytes = all.str();
  if (bytes.size() >= 5 && bytes.compare(0, 5, "solid") == 0) {
    if (stl_ascii(bytes, xyz)) return true;
  }
  return stl_binary(bytes, xyz);
}

namespace {

// appends "%.6g" of v (what std::ostream << float prints by default)
inline void put_num(std::string& s, float v) {
  char buf[32];
  int n = std::snprintf(buf, sizeof(buf), "%.*g", 6, double(v));
  s.append(buf, size_t(n));
}

std::string vtk_header(const char* kind, int nx, int ny, int nz, size_t n, const char* field,
                       const char* encoding = "ASCII") {
  std::string h = "# vtk DataFile Version 3.0\n";
  h += kind;
  h += "\n";
  h += encoding;
  h += "\nDATASET STRUCTURED_POINTS\nDIMENSIONS ";
  h += std::to_string(nx) + " " + std::to_string(ny) + " " + std::to_string(nz) + "\n";
  h += "ORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA " + std::to_string(n) + "\n";
  h += field;
  return h;
}

template <class Fmt>
bool write_parallel(const char* path, const std::string& header, size_t n, Fmt fmt) {
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size();
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 65536) nt = 1;
  const size_t chunk = 1 << 20;  // values per slice
  for (size_t base = 0; ok && base < n; base += chunk * nt) {
    std::vector<std::string> parts(nt);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
      const size_t b = base + t * chunk, e = std::min(n, b + chunk);
      if (b >= e) break;
      th.emplace_back([&, t, b, e] {
        std::string& s = parts[t];
        s.reserve((e - b) * 14);
        for (size_t i = b; i < e; ++i) fmt(s, i);
      });
    }
    for (auto& x : th) x.join();
    for (auto& p : parts)
      if (!p.empty() && std::fwrite(p.data(), 1, p.size(), f) != p.size()) ok = false;
  }
  if (std::fclose(f) != 0) ok = false;
  return ok;
}

// big-endian float32 of values[k * stride + i] for i in [0, n), k in [0, comps)
// interleaved per point, then a closing newline
bool write_binary(const char* path, const std::string& header, size_t n,
                  const float* const* comps, int ncomp) {
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size();
  const size_t chunk = size_t(1) << 20;
  std::vector<uint32_t> buf;
  for (size_t base = 0; ok && base < n; base += chunk) {
    const size_t e = std::min(n, base + chunk);
    buf.resize((e - base) * size_t(ncomp));
    size_t o = 0;
    for (size_t i = base; i < e; ++i)
      for (int k = 0; k < ncomp; ++k) {
        uint32_t u;
        std::memcpy(&u, comps[k] + i, 4);
        buf[o++] = __builtin_bswap32(u);
      }
    ok = std::fwrite(buf.data(), 4, buf.size(), f) == buf.size();
  }
  if (ok) ok = std::fputc('\n', f) != EOF;
  if (std::fclose(f) != 0) ok = false;
  return ok;
}

}  // namespace

bool vtk_scalar_binary(const char* path, const float* h, int nx, int ny, int nz) {
  const size_t n = size_t(nx) * ny * nz;
  const std::string hdr = vtk_header("scalar", nx, ny, nz, n,
                                     "SCALARS s float 1\nLOOKUP_TABLE default\n", "BINARY");
  const float* comps[1] = {h};
  return write_binary(path, hdr, n, comps, 1);
}

bool vtk_vector_binary(const char* path, const float* a, const float* b, const float* c, int nx,
                       int ny, int nz) {
  const size_t n = size_t(nx) * ny * nz;
  const std::string hdr = vtk_header("vector", nx, ny, nz, n, "VECTORS vec float\n", "BINARY");
  const float* comps[3] = {a, b, c};
  return write_binary(path, hdr, n, comps, 3);
}

bool vtk_scalar(const char* path, const float* h, int nx, int ny, int nz) {
  const size_t n = size_t(nx) * ny * nz;
  const std::string hdr =
      vtk_header("scalar", nx, ny, nz, n, "SCALARS s float 1\nLOOKUP_TABLE default\n");
  return write_parallel(path, hdr, n, [&](std::string& s, size_t i) {
    put_num(s, h[i]);
    s += '\n';
  });
}

bool vtk_vector(const char* path, const float* a, const float* b, const float* c, int nx, int ny,
                int nz) {
  const size_t n = size_t(nx) * ny * nz;
  const std::string hdr = vtk_header("vector", nx, ny, nz, n, "VECTORS vec float\n");
  return write_parallel(path, hdr, n, [&](std::string& s, size_t i) {
    put_num(s, a[i]);
    s += ' ';
    put_num(s, b[i]);
    s += ' ';
    put_num(s, c[i]);
    s += '\n';
  });
}

}  // namespace fs

// flb_io.cpp -- readback to VTK ImageData (.vti) files and the ParaView time
// series file, VTIWriter (src/io/writer.cpp:159-252): ASCII output in the
// reference's exact text layout, and a raw appended-binary variant for large
// grids.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/flb_core.h"

namespace {

// std::to_string(double): "%f".
std::string fixed6(double v) {
  char b[64];
  std::snprintf(b, sizeof b, "%f", v);
  return b;
}

// std::ostream << double with the default precision 6: "%g".
void put_g(std::string& out, double v) {
  char b[32];
  const int k = std::snprintf(b, sizeof b, "%g", v);
  out.append(b, size_t(k));
}

const char* type_name(int32_t t) { return t == FLB_VTI_UINT8 ? "UInt8" : "Float32"; }

double value(const flb_vti_field& f, size_t i) {
  if (f.type == FLB_VTI_UINT8) return double(static_cast<const uint8_t*>(f.data)[i]);
  return double(static_cast<const float*>(f.data)[i]);
}

}  // namespace

extern "C" int flb_write_vti(const char* path, uint32_t nx, uint32_t ny, double origin_x,
                             double origin_y, double spacing, const flb_vti_field* fields,
                             uint32_t nfields, int32_t binary) {
  if (!path || nx == 0 || ny == 0 || (nfields && !fields)) return FLB_ERR_ARGUMENT;
  for (uint32_t k = 0; k < nfields; ++k) {
    const flb_vti_field& f = fields[k];
    if (!f.name || !f.data || (f.components != 1 && f.components != 2) ||
        (f.type != FLB_VTI_UINT8 && f.type != FLB_VTI_FLOAT32) ||
        (f.type == FLB_VTI_UINT8 && f.components != 1))
      return FLB_ERR_ARGUMENT;
  }
  std::FILE* fp = std::fopen(path, "wb");
  if (!fp) return FLB_ERR_ARGUMENT;
  const size_t N = size_t(nx) * ny;
  const std::string ext = "0 " + std::to_string(nx - 1) + " 0 " + std::to_string(ny - 1) + " 0 0";
  std::string head = "<VTKFile type=\"ImageData\" version=\"1.0\"";
  if (binary) head += " byte_order=\"LittleEndian\" header_type=\"UInt64\"";
  head += ">\n<ImageData WholeExtent=\"" + ext + "\" Origin=\"" + fixed6(origin_x) + " " +
          fixed6(origin_y) + " " + fixed6(0.0) + "\" Spacing=\"" + fixed6(spacing) + " " +
          fixed6(spacing) + " " + fixed6(spacing) + "\" >\n<Piece Extent=\"" + ext + "\" >\n<PointData>\n";
  std::fputs(head.c_str(), fp);
  uint64_t offset = 0;
  std::string body;
  for (uint32_t k = 0; k < nfields; ++k) {
    const flb_vti_field& f = fields[k];
    body = std::string("<DataArray type=\"") + type_name(f.type) + "\" Name=\"" + f.name +
           "\" NumberOfComponents=\"" + std::to_string(f.components) + "\" format=\"" +
           (binary ? "appended\" offset=\"" + std::to_string(offset) + "\"" : std::string("ascii\"")) + ">\n";
    if (binary) {
      body += "</DataArray>\n";
      const uint64_t bytes = N * f.components * (f.type == FLB_VTI_UINT8 ? 1 : 4);
      offset += sizeof(uint64_t) + bytes;
    } else {
      body.reserve(body.size() + N * 12 * f.components);
      const double ysign = f.flip_y ? -1.0 : 1.0;
      for (size_t i = 0; i < N; ++i) {
        if (f.components == 1) {
          put_g(body, value(f, i) * f.scale);
        } else {
          put_g(body, value(f, i) * f.scale);
          body += ' ';
          put_g(body, value(f, N + i) * f.scale * ysign);
        }
        body += '\n';
      }
      body += "</DataArray>\n";
    }
    std::fwrite(body.data(), 1, body.size(), fp);
  }
  std::string tail = "</PointData>\n<CellData>\n</CellData>\n</Piece>\n</ImageData>\n";
  if (binary) {
    tail += "<AppendedData encoding=\"raw\">\n_";
    std::fputs(tail.c_str(), fp);
    for (uint32_t k = 0; k < nfields; ++k) {
      const flb_vti_field& f = fields[k];
      const uint64_t bytes = N * f.components * (f.type == FLB_VTI_UINT8 ? 1 : 4);
      std::fwrite(&bytes, sizeof bytes, 1, fp);
      if (f.type == FLB_VTI_UINT8 || (f.components == 1 && f.scale == 1.0)) {
        std::fwrite(f.data, 1, size_t(bytes), fp);
      } else {  // interleave components, apply scale and y sign
        std::vector<float> buf(N * f.components);
        const float* d = static_cast<const float*>(f.data);
        const double ysign = f.flip_y ? -1.0 : 1.0;
        for (size_t i = 0; i < N; ++i) {
          if (f.components == 1) buf[i] = float(d[i] * f.scale);
          else {
            buf[2 * i] = float(d[i] * f.scale);
            buf[2 * i + 1] = float(d[N + i] * f.scale * ysign);
          }
        }
        std::fwrite(buf.data(), 4, buf.size(), fp);
      }
    }
    tail = "\n</AppendedData>\n";
  }
  tail += "</VTKFile>\n";
  std::fputs(tail.c_str(), fp);
  const bool ok = std::ferror(fp) == 0;
  std::fclose(fp);
  return ok ? FLB_OK : FLB_ERR_ARGUMENT;
}

// Writer::createTimeSeriesFile (src/io/writer.cpp:159-181).
extern "C" int flb_write_vti_series(const char* path, uint32_t count, double time_interval) {
  if (!path) return FLB_ERR_ARGUMENT;
  std::FILE* fp = std::fopen(path, "wb");
  if (!fp) return FLB_ERR_ARGUMENT;
  std::string s = "{\n\"file-series-version\" : \"1.0\",\n\"files\" : [\n";
  for (uint32_t i = 0; i < count; ++i)
    s += "{\"name\" : \"result_" + std::to_string(i) + ".vti\", \"time\" : " +
         std::to_string(i * time_interval) + "},\n";
  s += "]\n}";
  std::fputs(s.c_str(), fp);
  const bool ok = std::ferror(fp) == 0;
  std::fclose(fp);
  return ok ? FLB_OK : FLB_ERR_ARGUMENT;
}

// Readback of a handle's fields into a .vti: Flags, U (scaled), rho (scaled)
// and phi for free-surface handles, as the run loops save them
// (src/cuda/runCuda.cu:483-487, 518-523).
extern "C" int flb_save_vti(flb_handle* h, const char* path, double origin_x, double origin_y,
                            double spacing, double velocity_scale, double density_scale,
                            int32_t binary) {
  if (!h || !path) return FLB_ERR_ARGUMENT;
  uint32_t nx = 0, ny = 0;
  int32_t problem = 0;
  int rc0 = flb_get_shape(h, &nx, &ny, &problem);
  if (rc0) return rc0;
  const bool free_surface = problem == FLB_FREE_SURFACE;
  const size_t N = size_t(nx) * ny;
  std::vector<float> u(2 * N), rho(N), phi(free_surface ? N : 0);
  std::vector<uint8_t> flags(N);
  int rc = flb_read(h, u.data(), rho.data(), flags.data(), free_surface ? phi.data() : nullptr);
  if (rc) return rc;
  flb_vti_field f[4] = {{"Flags", FLB_VTI_UINT8, 1, flags.data(), 1.0, 0},
                        {"U", FLB_VTI_FLOAT32, 2, u.data(), velocity_scale, 0},
                        {"rho", FLB_VTI_FLOAT32, 1, rho.data(), density_scale, 0},
                        {"phi", FLB_VTI_FLOAT32, 1, phi.data(), 1.0, 0}};
  return flb_write_vti(path, nx, ny, origin_x, origin_y, spacing, f, free_surface ? 4u : 3u, binary);
}

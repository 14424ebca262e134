// Field output writers (pf_io.hpp) and the context-free C-ABI entry point
// pf_write_vtk.
#include "pf_io.hpp"

#include <cinttypes>
#include <cstring>

#include "../../include/phasefieldlbm.h"

namespace pfio {

std::string step_suffix(uint64_t step) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "%06" PRIu64, step);
    return buf;
}

std::string join(const std::string& dir, const std::string& name) {
    if (dir.empty()) return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

bool write_bin(const std::string& path, const float* data, size_t n) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, sizeof(float), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

VtkWriter::~VtkWriter() {
    if (f_) std::fclose(f_);
}

bool VtkWriter::open(const std::string& path, int kind, int nx, int ny, int nz,
                     const std::vector<std::string>& names) {
    if (f_ || (kind != kVti && kind != kVtr) || names.empty()) return false;
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) return false;
    kind_ = kind;
    nx_ = nx, ny_ = ny, nz_ = nz;
    const uint64_t nbytes = (uint64_t)nx * ny * nz * sizeof(float);
    wide_ = nbytes >= ((uint64_t)1 << 32);
    const uint64_t hdr = wide_ ? 8 : 4;
    const char* type = kind == kVti ? "ImageData" : "RectilinearGrid";
    std::string x;
    x += "<?xml version=\"1.0\"?>\n";
    x += std::string("<VTKFile type=\"") + type + "\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"" +
         (wide_ ? "UInt64" : "UInt32") + "\">\n";
    char ext[128];
    std::snprintf(ext, sizeof ext, "0 %d 0 %d 0 %d", nx - 1, ny - 1, nz - 1);
    if (kind == kVti)
        x += std::string("  <ImageData WholeExtent=\"") + ext + "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    else
        x += std::string("  <RectilinearGrid WholeExtent=\"") + ext + "\">\n";
    x += std::string("    <Piece Extent=\"") + ext + "\">\n";
    x += "      <PointData Scalars=\"" + names.front() + "\">\n";
    uint64_t off = 0;
    for (const std::string& n : names) {
        x += "        <DataArray type=\"Float32\" Name=\"" + n +
             "\" NumberOfComponents=\"1\" format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
        off += hdr + nbytes;
    }
    x += "      </PointData>\n      <CellData/>\n";
    if (kind == kVtr) {
        // cell-centre coordinates 0..n-1 along each axis, unit spacing as in the .vti
        x += "      <Coordinates>\n";
        const int len[3] = {nx, ny, nz};
        const char* axis[3] = {"x", "y", "z"};
        for (int a = 0; a < 3; ++a) {
            x += std::string("        <DataArray type=\"Float32\" Name=\"") + axis[a] +
                 "\" NumberOfComponents=\"1\" format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
            off += hdr + (uint64_t)len[a] * sizeof(float);
        }
        x += "      </Coordinates>\n";
    }
    x += "    </Piece>\n";
    x += kind == kVti ? "  </ImageData>\n" : "  </RectilinearGrid>\n";
    x += "  <AppendedData encoding=\"raw\">\n_";
    expected_ = names.size();
    written_ = 0;
    return std::fwrite(x.data(), 1, x.size(), f_) == x.size();
}

bool VtkWriter::put_bytes(const void* data, uint64_t nbytes) {
    bool ok;
    if (wide_) {
        ok = std::fwrite(&nbytes, 8, 1, f_) == 1;
    } else {
        const uint32_t n32 = (uint32_t)nbytes;
        ok = std::fwrite(&n32, 4, 1, f_) == 1;
    }
    return ok && std::fwrite(data, 1, (size_t)nbytes, f_) == (size_t)nbytes;
}

bool VtkWriter::put(const float* data) {
    if (!f_ || written_ >= expected_) return false;
    ++written_;
    return put_bytes(data, (uint64_t)nx_ * ny_ * nz_ * sizeof(float));
}

bool VtkWriter::close() {
    if (!f_) return false;
    bool ok = written_ == expected_;
    if (kind_ == kVtr) {
        const int len[3] = {nx_, ny_, nz_};
        for (int a = 0; a < 3 && ok; ++a) {
            std::vector<float> c((size_t)len[a]);
            for (int i = 0; i < len[a]; ++i) c[(size_t)i] = (float)i;
            ok = put_bytes(c.data(), (uint64_t)len[a] * sizeof(float));
        }
    }
    const char tail[] = "\n  </AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail, 1, sizeof tail - 1, f_) == sizeof tail - 1;
    ok = std::fclose(f_) == 0 && ok;
    f_ = nullptr;
    return ok;
}

}  // namespace pfio

extern "C" int pf_write_vtk(const char* path, int32_t format, int32_t nx, int32_t ny, int32_t nz,
                            const char* const* names, const float* const* arrays, int32_t narrays) {
    if (!path || !names || !arrays || narrays < 1 || nx < 1 || ny < 1 || nz < 1) return PF_ERR_CONFIG;
    if (format != PF_OUT_VTI && format != PF_OUT_VTR) return PF_ERR_CONFIG;
    std::vector<std::string> nm;
    for (int32_t i = 0; i < narrays; ++i) {
        if (!names[i] || !arrays[i] || !names[i][0] || std::strpbrk(names[i], "\"<>&")) return PF_ERR_CONFIG;
        nm.emplace_back(names[i]);
    }
    pfio::VtkWriter w;
    if (!w.open(path, format == PF_OUT_VTI ? pfio::kVti : pfio::kVtr, nx, ny, nz, nm)) return PF_ERR_IO;
    for (int32_t i = 0; i < narrays; ++i)
        if (!w.put(arrays[i])) return PF_ERR_IO;
    return w.close() ? PF_OK : PF_ERR_IO;
}

// NPZ output of one analysed pair (piv-io/src/write.rs:113-165, NpzWriter): a zip archive of
// .npy arrays, deflate-compressed (zlib, raw deflate streams) as ndarray-npy's
// NpzWriter::new_compressed writes them.
//
// Keys and types: x, y (f32 [nx], [ny], grid centres plus the ROI offset), dx, dy (f32
// [ny, nx]), u, v (= dx, dy times f32(1 / dt)), u_mps, v_mps (u, v times f32(m_per_px), only
// with a scale), peak_ratio (f32), flags (u8), frames (u64 [2]), t (f64 scalar).
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "piv_internal.hpp"

using namespace pivi;

namespace {

struct NpyArray {
    std::string name;
    std::string descr;          // '<f4', '|u1', '<u8', '<f8'
    std::vector<size_t> shape;  // empty: scalar
    const void *data;
    size_t bytes;
};

std::string npy_bytes(const NpyArray &a)
{
    std::string shape = "(";
    for (size_t i = 0; i < a.shape.size(); ++i) {
        shape += std::to_string(a.shape[i]);
        shape += (a.shape.size() == 1 || i + 1 < a.shape.size()) ? "," : "";
        if (i + 1 < a.shape.size()) shape += " ";
    }
    shape += ")";
    std::string dict = "{'descr': '" + a.descr + "', 'fortran_order': False, 'shape': " + shape + ", }";
    // format 1.0: magic, version, u16 header length; header padded with spaces to a multiple
    // of 64 bytes (including the 10-byte preamble) and ended by a newline
    size_t total = 10 + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict.append(pad, ' ');
    dict += '\n';
    std::string out;
    out += "\x93NUMPY";
    out += (char)1;
    out += (char)0;
    const uint16_t hl = (uint16_t)dict.size();
    out += (char)(hl & 0xff);
    out += (char)(hl >> 8);
    out += dict;
    out.append(static_cast<const char *>(a.data), a.bytes);
    return out;
}

void put16(std::string &s, uint16_t v)
{
    s += (char)(v & 0xff);
    s += (char)(v >> 8);
}

void put32(std::string &s, uint32_t v)
{
    for (int i = 0; i < 4; ++i) s += (char)((v >> (8 * i)) & 0xff);
}

// Raw deflate with zlib's run-length strategy: runs (flags, coordinates) still collapse and
// float fields get their Huffman coding, at over twice the speed of the default LZ77 search,
// which finds almost nothing in float fields (same size to 0.1 %).  Any inflater reads it.
bool deflate_raw(const std::string &in, std::string &out)
{
    z_stream zs{};
    if (deflateInit2(&zs, 6, Z_DEFLATED, -15, 8, Z_RLE) != Z_OK) return false;
    out.resize(deflateBound(&zs, (uLong)in.size()) + 16);
    zs.next_in = reinterpret_cast<Bytef *>(const_cast<char *>(in.data()));
    zs.avail_in = (uInt)in.size();
    zs.next_out = reinterpret_cast<Bytef *>(&out[0]);
    zs.avail_out = (uInt)out.size();
    const int r = deflate(&zs, Z_FINISH);
    out.resize(zs.total_out);
    deflateEnd(&zs);
    return r == Z_STREAM_END;
}

// Writes a zip archive with one deflated member per array (no zip64: members < 4 GiB).
piv_status write_npz(const std::string &path, const std::vector<NpyArray> &arrays)
{
    std::string zip, central;
    uint16_t count = 0;
    for (const NpyArray &a : arrays) {
        const std::string raw = npy_bytes(a);
        std::string comp;
        if (!deflate_raw(raw, comp)) return fail(PIV_ERR_INVALID_ARGUMENT, "deflate failed");
        const uint32_t crc = (uint32_t)crc32(0L, reinterpret_cast<const Bytef *>(raw.data()), (uInt)raw.size());
        const std::string fname = a.name + ".npy";
        const uint32_t offset = (uint32_t)zip.size();
        // local file header
        put32(zip, 0x04034b50);
        put16(zip, 20);       // version needed
        put16(zip, 0);        // flags
        put16(zip, 8);        // deflate
        put16(zip, 0);        // time
        put16(zip, 0x21);     // date (1980-01-01)
        put32(zip, crc);
        put32(zip, (uint32_t)comp.size());
        put32(zip, (uint32_t)raw.size());
        put16(zip, (uint16_t)fname.size());
        put16(zip, 0);
        zip += fname;
        zip += comp;
        // central directory entry
        put32(central, 0x02014b50);
        put16(central, 20);   // version made by
        put16(central, 20);
        put16(central, 0);
        put16(central, 8);
        put16(central, 0);
        put16(central, 0x21);
        put32(central, crc);
        put32(central, (uint32_t)comp.size());
        put32(central, (uint32_t)raw.size());
        put16(central, (uint16_t)fname.size());
        put16(central, 0);    // extra
        put16(central, 0);    // comment
        put16(central, 0);    // disk
        put16(central, 0);    // internal attributes
        put32(central, 0);    // external attributes
        put32(central, offset);
        central += fname;
        ++count;
    }
    const uint32_t cd_off = (uint32_t)zip.size();
    zip += central;
    put32(zip, 0x06054b50);
    put16(zip, 0);
    put16(zip, 0);
    put16(zip, count);
    put16(zip, count);
    put32(zip, (uint32_t)central.size());
    put32(zip, cd_off);
    put16(zip, 0);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(PIV_ERR_INVALID_ARGUMENT, "cannot create %s", path.c_str());
    const size_t w = std::fwrite(zip.data(), 1, zip.size(), f);
    const int c = std::fclose(f);
    if (w != zip.size() || c != 0) return fail(PIV_ERR_INVALID_ARGUMENT, "writing %s failed", path.c_str());
    return PIV_OK;
}

}  // namespace

extern "C" {

piv_status piv_write_pair_npz(const char *dir, uint64_t frame_a, uint64_t frame_b, double time_s,
                              double dt_s, double m_per_px, uint32_t roi_x, uint32_t roi_y,
                              const piv_grid *grid, const float *dx, const float *dy,
                              const float *peak_ratio, const uint8_t *flags)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!dir || !grid || !dx || !dy || !peak_ratio || !flags)
            return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        if (!(dt_s > 0.0)) return fail(PIV_ERR_INVALID_ARGUMENT, "dt must be positive");
        const size_t nx = grid->nx, ny = grid->ny, n = nx * ny;
        // grid.rs:60-74 centres, plus the ROI offset (write.rs grid_coords)
        std::vector<float> xs(nx), ys(ny);
        const float half = ((float)grid->window - 1.0f) / 2.0f;
        for (size_t i = 0; i < nx; ++i)
            xs[i] = ((float)(grid->x0 + i * grid->step) + half) + (float)roi_x;
        for (size_t i = 0; i < ny; ++i)
            ys[i] = ((float)(grid->y0 + i * grid->step) + half) + (float)roi_y;
        // write.rs derive(): u = dx * (1 / dt) as f32, physical = u * (m_per_px as f32)
        const float inv_dt = (float)(1.0 / dt_s);
        std::vector<float> u(n), v(n), um, vm;
        for (size_t i = 0; i < n; ++i) {
            u[i] = dx[i] * inv_dt;
            v[i] = dy[i] * inv_dt;
        }
        const bool physical = m_per_px > 0.0;
        if (physical) {
            const float k = (float)m_per_px;
            um.resize(n);
            vm.resize(n);
            for (size_t i = 0; i < n; ++i) {
                um[i] = u[i] * k;
                vm[i] = v[i] * k;
            }
        }
        const uint64_t frames[2] = {frame_a, frame_b};
        std::vector<NpyArray> arrays;
        arrays.push_back({"x", "<f4", {nx}, xs.data(), nx * 4});
        arrays.push_back({"y", "<f4", {ny}, ys.data(), ny * 4});
        arrays.push_back({"dx", "<f4", {ny, nx}, dx, n * 4});
        arrays.push_back({"dy", "<f4", {ny, nx}, dy, n * 4});
        arrays.push_back({"u", "<f4", {ny, nx}, u.data(), n * 4});
        arrays.push_back({"v", "<f4", {ny, nx}, v.data(), n * 4});
        if (physical) {
            arrays.push_back({"u_mps", "<f4", {ny, nx}, um.data(), n * 4});
            arrays.push_back({"v_mps", "<f4", {ny, nx}, vm.data(), n * 4});
        }
        arrays.push_back({"peak_ratio", "<f4", {ny, nx}, peak_ratio, n * 4});
        arrays.push_back({"flags", "|u1", {ny, nx}, flags, n});
        arrays.push_back({"frames", "<u8", {2}, frames, 16});
        arrays.push_back({"t", "<f8", {}, &time_s, 8});
        char name[64];
        std::snprintf(name, sizeof name, "/pair_%06llu_%06llu.npz", (unsigned long long)frame_a,
                      (unsigned long long)frame_b);
        return write_npz(std::string(dir) + name, arrays);
    });
}

piv_status piv_write_stats_npz(const char *dir, uint32_t npairs, const double *t,
                               const piv_flow_stats *stats)
{
    return pivi::guard(__func__, [&]() -> piv_status {
        if (!dir || (npairs && (!t || !stats))) return fail(PIV_ERR_INVALID_ARGUMENT, "null argument");
        // flow.rs:133-145 write_stats_npz / :147-156 columns: t (f64), then the f32 columns
        const size_t n = npairs;
        std::vector<float> col[5];
        for (auto &c : col) c.resize(n);
        for (size_t i = 0; i < n; ++i) {
            col[0][i] = stats[i].vorticity_mean;
            col[1][i] = stats[i].vorticity_abs_mean;
            col[2][i] = stats[i].ke_mean;
            col[3][i] = stats[i].ke_sum;
            col[4][i] = stats[i].coverage;
        }
        static const char *names[5] = {"vorticity_mean", "vorticity_abs_mean", "ke_mean", "ke_sum",
                                       "coverage"};
        std::vector<NpyArray> arrays;
        arrays.push_back({"t", "<f8", {n}, t, n * 8});
        for (int c = 0; c < 5; ++c) arrays.push_back({names[c], "<f4", {n}, col[c].data(), n * 4});
        return write_npz(std::string(dir) + "/stats.npz", arrays);
    });
}

}  // extern "C"

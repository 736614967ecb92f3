// qk_paths_io.cpp -- writer and reader of the ".paths" file (see qk_paths_io.hpp).
#include "qk_paths_io.hpp"

#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <stdexcept>

#include "qk_engine.hpp"

namespace quantkos::Engine {

namespace {
constexpr size_t HEADER_BYTES = 40;

void put_header(char *h, int precision_bytes, int n_sims, int n_steps, double dt, double scale) {
    std::memset(h, 0, HEADER_BYTES);
    std::memcpy(h, "QUANTKOS", 8);
    const int32_t ints[4] = {1, precision_bytes, n_sims, n_steps};
    std::memcpy(h + 8, ints, sizeof(ints));
    if (precision_bytes == 4) {
        const float f = (float)dt;
        std::memcpy(h + 24, &f, 4);
    } else {
        std::memcpy(h + 24, &dt, 8);
    }
    std::memcpy(h + 32, &scale, 8);
}
}  // namespace

std::string write_paths_file(const Config::UInputs &cfg, bool stored) {
    const std::filesystem::path file = std::filesystem::path(cfg.output.out_dir) / cfg.output.filename_paths_out;
    std::ofstream out(file, std::ios::out | std::ios::binary);
    if (!out) throw std::runtime_error("[BinWriter Error] Failed to open binary output file: " + file.string());
    const qk_contract c = to_contract(cfg);
    qk_ctx *ctx = device_context();
    if (!stored) {
        const qk_payoff p = to_payoff(cfg);
        qk_stats s{};
        if (qk_price_store_paths(ctx, &c, &p, &s) != QK_OK)
            throw std::runtime_error(std::string("qk_price_store_paths failed: ") + qk_last_error(ctx));
    }
    const bool fp32 = c.precision == QK_FP32;
    const size_t rs = fp32 ? 4 : 8;
    char header[HEADER_BYTES];
    put_header(header, (int)rs, cfg.mc.N_Paths, cfg.time.N_time_steps, cfg.time.dt, cfg.get_price_scaling_factor());
    out.write(header, HEADER_BYTES);
    // The device hands back row-major [path][step] windows of ~64 MiB in the
    // run's precision: the payload is written as it arrives.
    const int64_t N = cfg.mc.N_Paths, T = cfg.time.N_time_steps;
    const int64_t chunk = std::max<int64_t>(1, (int64_t)(64u << 20) / (T * (int64_t)rs));
    std::vector<char> buf;
    for (int64_t p0 = 0; p0 < N; p0 += chunk) {
        const int64_t n = std::min(chunk, N - p0);
        buf.resize((size_t)(n * T) * rs);
        if (qk_copy_stored_paths(ctx, p0, n, buf.data()) != QK_OK)
            throw std::runtime_error(std::string("qk_copy_stored_paths failed: ") + qk_last_error(ctx));
        out.write(buf.data(), (std::streamsize)buf.size());
    }
    if (!out) throw std::runtime_error("[BinWriter Error] write failed: " + file.string());
    return file.string();
}

PathsFile PathsFile::open(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("[BinReader Error] Cannot open " + path);
    char h[HEADER_BYTES];
    in.read(h, HEADER_BYTES);
    if (!in || std::memcmp(h, "QUANTKOS", 8) != 0) throw std::runtime_error("[BinReader Error] Not a QuantKos file: " + path);
    PathsFile f;
    int32_t ints[4];
    std::memcpy(ints, h + 8, sizeof(ints));
    f.version = ints[0];
    f.byte_precision = ints[1];
    f.total_n_sims = ints[2];
    f.n_time_steps = ints[3];
    if (f.byte_precision == 4) {
        float d;
        std::memcpy(&d, h + 24, 4);
        f.dt = d;
    } else if (f.byte_precision == 8) {
        std::memcpy(&f.dt, h + 24, 8);
    } else {
        throw std::runtime_error("[BinReader Error] Unsupported precision in " + path);
    }
    std::memcpy(&f.scale_factor, h + 32, 8);
    f.path = path;
    return f;
}

std::vector<double> PathsFile::read_prices_at_target_time(double t) const {
    const int col = (int)std::lround(t / dt) - 1;
    if (col < 0 || col >= n_time_steps)
        throw std::runtime_error("[BinReader Error] Requested time " + std::to_string(t) +
                                 " is outside the simulated time grid bounds.");
    if (std::abs(t - (col + 1) * dt) > 1e-10 * std::max(1.0, std::abs(t)) + 1e-6 * dt)
        throw std::runtime_error("Requested time " + std::to_string(t) + " does not align with the simulation grid.");
    std::ifstream in(path, std::ios::binary);
    std::vector<double> out((size_t)total_n_sims);
    for (int i = 0; i < total_n_sims; ++i) {
        in.seekg((std::streamoff)(HEADER_BYTES + ((int64_t)i * n_time_steps + col) * byte_precision));
        if (byte_precision == 4) {
            float v;
            in.read(reinterpret_cast<char *>(&v), 4);
            out[(size_t)i] = v * scale_factor;
        } else {
            double v;
            in.read(reinterpret_cast<char *>(&v), 8);
            out[(size_t)i] = v * scale_factor;
        }
    }
    return out;
}

}  // namespace quantkos::Engine

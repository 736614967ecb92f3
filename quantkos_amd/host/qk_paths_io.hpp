// qk_paths_io.hpp -- the binary ".paths" file of simulated spot paths.
//
// Same layout as the reference's BinWriter/BinReader (include/IO/IOBinary.hpp:
// header :72-81, writer :139-177, strided reader :297-344): a 40-byte header
//   char key[8] = "QUANTKOS"; int32 version = 1; int32 byte_precision (4|8);
//   int32 total_n_sims; int32 n_time_steps; Real dt (+4 pad when FP32);
//   double scale_factor
// followed by the row-major [path][step] matrix of normalised spot prices in
// the run's precision (multiply by scale_factor for real prices).
#pragma once

#include <string>
#include <vector>

#include "qk_config.hpp"

namespace quantkos::Engine {

// Writes out_dir/Name_Paths_Out_File from the path set the pricing call left
// in the device ctx's path store (stored = true: qk_price_store_paths or
// qk_price_american* just ran on `scaled`), or prices once more with the
// store when stored = false.  `scaled` is the configuration after
// apply_price_scaling.  Returns the file path.
std::string write_paths_file(const Config::UInputs &scaled, bool stored);

struct PathsFile {
    int version = 0, byte_precision = 0, total_n_sims = 0, n_time_steps = 0;
    double dt = 0.0, scale_factor = 1.0;
    std::string path;
    static PathsFile open(const std::string &path);
    // Real prices (scaled back) of every path at time t = (k+1) dt (BinReader::read_prices_at_target_time).
    std::vector<double> read_prices_at_target_time(double t) const;
};

}  // namespace quantkos::Engine

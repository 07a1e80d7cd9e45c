// cli.cpp -- command-line front end: pgure-svt <paramfile>
// (src/PGURE-SVT.cpp:18-246), a second caller of the same entry point.
//
// Reads frames start_frame..end_frame (1-based) of <stem>.tif, optionally
// removes hot pixels, denoises them with pgsvt_denoise as uint16 data, and
// writes <stem>-CLEANED.tif (16-bit), with the reference's parameter names,
// defaults and messages.  Deliberately kept from the reference:
//   * the parameter-file syntax of ParseParameters (src/utils.hpp:55-94):
//     "name : value", '#' comments, the value is the concatenation of the
//     tokens after the separator up to one holding '#';
//   * `tolerance` is read but not used (the reference parses it into a
//     shadowing local, src/PGURE-SVT.cpp:93-99), so tol stays 1e-7;
//   * `hot_pixel` is read as an integer (std::stoi);
//   * max_iter defaults to 1000 (500 in the Python API).
// Deviations: values above 65535 are clamped when converting the result to
// uint16 (undefined in the reference); 8-bit input is widened pixel by pixel
// (the reference reads 8-bit scanlines into a uint16 buffer); with
// optimize_pgure on, the lambda search starts from the window mean (the
// reference's start of 0 makes NLopt reject a zero initial step).
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pgsvt.h"

namespace {

using Clock = std::chrono::steady_clock;

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// "name : value" lines (src/utils.hpp:55-90)
std::map<std::string, std::string> parse_parameters(std::istream& in) {
  std::map<std::string, std::string> opts;
  for (std::string line; std::getline(in, line);) {
    std::istringstream iss(line);
    std::string id, sep, val, tok;
    if (!(iss >> id) || id[0] == '#') continue;
    // a value follows when the separator is ":" or when more text follows it
    if (!(iss >> sep) || sep == ":" || iss.get() != EOF) {
      while (iss >> tok) {
        if (tok.find('#') != std::string::npos) break;
        val += tok;
      }
    }
    opts[id] = val;
  }
  return opts;
}

bool str_to_bool(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  return s == "1" || s == "true";
}

int error(const std::string& msg) {
  std::cerr << "**ERROR**\n" << msg << "\n";
  return -1;
}

int run(int argc, char** argv) {
  std::cout << "PGURE-SVT Denoising (" << pgsvt_version() << ")\n";
  if (argc != 2) {
    std::cout << "  Usage: pgure-svt paramfile\n";
    return -1;
  }
  std::ifstream pf(argv[1]);
  auto opts = parse_parameters(pf);
  if (!opts.count("filename") || !opts.count("start_frame") || !opts.count("end_frame"))
    return error("Required parameters not specified\nYou must specify 'filename', 'start_frame' and 'end_frame'");

  auto get_i = [&](const char* k, long d) { return opts.count(k) ? std::stol(opts.at(k)) : d; };
  auto get_d = [&](const char* k, double d) { return opts.count(k) ? std::stod(opts.at(k)) : d; };
  auto get_b = [&](const char* k, bool d) { return opts.count(k) ? str_to_bool(opts.at(k)) : d; };

  const std::string filename = opts.at("filename");
  const std::string stem = filename.substr(0, filename.find_last_of('.'));
  const uint32_t start = (uint32_t)std::stoi(opts.at("start_frame"));
  const uint32_t end = (uint32_t)std::stoi(opts.at("end_frame"));
  if (start < 1 || end < start) return error("Frame range start_frame..end_frame is empty");
  const uint32_t n_images = end - start + 1;

  pgsvt_params p;
  pgsvt_default_params(&p);
  p.patch_size = (uint32_t)get_i("patch_size", 4);
  p.trajectory_length = (uint32_t)get_i("trajectory_length", 15);
  p.motion_window = (uint32_t)get_i("motion_neighbourhood", 7);
  p.motion_filter = (int64_t)(uint32_t)get_i("median_filter", 5);
  p.patch_overlap = (uint32_t)get_i("patch_overlap", 1);
  p.noise_method = (uint32_t)get_i("noise_method", 4);
  p.max_iter = (uint32_t)get_i("max_iter", 1000);
  p.n_jobs = get_i("n_jobs", -1);
  p.noise_alpha = get_d("noise_alpha", -1.);
  p.noise_mu = get_d("noise_mu", -1.);
  p.noise_sigma = get_d("noise_sigma", -1.);
  const double hot_pixel = opts.count("hot_pixel") ? (double)std::stoi(opts.at("hot_pixel")) : -1.0;
  p.random_seed = get_i("random_seed", -1);
  p.exponential_weighting = get_b("exponential_weighting", true);
  const bool normalize = get_b("normalize", false);
  p.motion_estimation = get_b("motion_estimation", true);
  p.optimize_pgure = get_b("optimize_pgure", true);
  // With the search on, the reference passes lambda = 0, whose initial step
  // sqrt(0) NLopt rejects (the run aborts); start from the window mean
  // instead, as the Python API does.
  p.lambda1 = -1.0;
  if (!p.optimize_pgure) {
    if (!opts.count("lambda")) return error("PGURE optimization is turned OFF but no lambda specified in parameter file");
    p.lambda1 = std::stod(opts.at("lambda"));
  }
  p.tol = 1e-7;  // `tolerance` is not applied, as in the reference

  auto t0 = Clock::now();
  const std::string in_name = stem + ".tif";
  if (!std::ifstream(in_name)) return error("File " + in_name + " not found");
  uint32_t W = 0, H = 0, bits = 0, pages = 0;
  if (pgsvt_tiff_info(in_name.c_str(), &W, &H, &bits, &pages)) return error(pgsvt_last_error());
  if (W != H) return error("Frame dimensions are not square, got " + std::to_string(W) + "x" + std::to_string(H));
  if (bits != 8 && bits != 16) return error("Images must be 8-bit or 16-bit, got " + std::to_string(bits) + "-bit depth");
  const uint32_t avail = (pages >= start) ? std::min(pages - (start - 1), n_images) : 0;
  if (avail < n_images)
    return error("Sequence only has " + std::to_string(avail) + " frames, expected " + std::to_string(n_images));
  std::vector<uint16_t> seq((size_t)W * H * n_images);
  if (pgsvt_tiff_read_u16(in_name.c_str(), start - 1, n_images, seq.data())) return error(pgsvt_last_error());
  auto t1 = Clock::now();
  std::cout << std::fixed << std::setprecision(4) << "TIFF import:    " << std::setw(10) << seconds(t0, t1)
            << " seconds\n";

  if (hot_pixel >= 0.0) {
    t0 = Clock::now();
    if (pgsvt_hotpixel_filter(seq.data(), H, W, n_images, hot_pixel)) return error(pgsvt_last_error());
    t1 = Clock::now();
    std::cout << "Outlier filter: " << std::setw(10) << seconds(t0, t1) << " seconds\n";
  }

  t0 = Clock::now();
  std::vector<double> Y(seq.size()), est((size_t)n_images * 4);
  if (pgsvt_denoise(seq.data(), PGSVT_U16, H, W, n_images, &p, Y.data(), est.data()))
    return error(pgsvt_last_error());
  t1 = Clock::now();
  std::cout << "PGURE-SVT:      " << std::setw(10) << seconds(t0, t1) << " seconds\n";

  t0 = Clock::now();
  if (normalize) {  // to [0, 65535]
    const auto mm = std::minmax_element(Y.begin(), Y.end());
    const double lo = *mm.first, hi = *mm.second;
    for (double& v : Y) v = 65535 * (v - lo) / (hi - lo);
  }
  std::vector<uint16_t> out(Y.size());
  for (size_t i = 0; i < Y.size(); ++i) {
    const double v = Y[i];
    out[i] = (!(v > 0.0)) ? 0 : (v >= 65535.0 ? 65535 : (uint16_t)v);  // truncation, negatives to 0
  }
  const std::string out_name = stem + "-CLEANED.tif";
  if (pgsvt_tiff_write_u16(out_name.c_str(), out.data(), W, H, n_images))
    return error("File " + out_name + " could not be written");
  t1 = Clock::now();
  std::cout << "TIFF export:    " << std::setw(10) << seconds(t0, t1) << " seconds\n\n";
  std::cerr << "Output file:    " << out_name << "\n";
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {  // std::stoi / std::stod on a malformed value
    return error(std::string("Invalid parameter value (") + e.what() + ")");
  }
}

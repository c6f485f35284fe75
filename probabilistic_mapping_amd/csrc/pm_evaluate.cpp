// pm_evaluate: a file-based stand-in for the reference's ROS node (evaluate.cc).
//
//   pm_evaluate IMAGES POSES OUT_DIR [CASE] [--fx F] [--fy F] [--cx F] [--cy F]
//               [--agg-num N] [--interval N] [--planes N] [--dep-sample F] [--plan-only]
//               [--no-frame-ring]
//
// IMAGES: lines "timestamp path" (TUM rgb.txt layout; '#' starts a comment; a relative path is
// taken from IMAGES' directory), each an 8-bit binary PGM (P5).  POSES: lines
// "timestamp tx ty tz qx qy qz qw" (TUM trajectory layout), camera -> world.  The messages are
// replayed in timestamp order (a pose before an image of the same stamp, as its topic would
// have to arrive for the step to run) and one processImg step (pm::FrameDriver) follows every
// message, as the node's 1 kHz loop would.  Each depth (the node's 32FC1 /depth message,
// evaluate.cc:262-264) is written to OUT_DIR/<timestamp>.npy (float32 [H][W], little-endian)
// and listed in OUT_DIR/depths.txt as "timestamp file reference-image source-images...".
// CASE is the reference's command-line mode (evaluate.cc:313-319).  --plan-only prints the
// steps without running the mapper (no GPU needed).  --no-frame-ring uploads every frame at
// every use, as the reference does, instead of once through the frame ring (same depths).
#include <algorithm>
#include <cctype>
#include <stdexcept>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "pm/frame_driver.h"

namespace {

// "1305031102.175304" -> nanoseconds, exactly
bool parse_stamp(const std::string& s, int64_t* ns) {
  const size_t dot = s.find('.');
  const std::string whole = s.substr(0, dot);
  std::string frac = dot == std::string::npos ? "" : s.substr(dot + 1);
  if (whole.empty() || frac.size() > 9) return false;
  for (char c : whole + frac)
    if (c < '0' || c > '9') return false;
  frac.resize(9, '0');
  *ns = std::stoll(whole) * 1000000000LL + std::stoll(frac);
  return true;
}

std::string stamp_text(int64_t ns) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%lld.%09lld", (long long)(ns / 1000000000LL),
                (long long)(ns % 1000000000LL));
  return buf;
}

// next whitespace-separated PGM header token, skipping '#' comments
bool pgm_token(std::istream& in, std::string& tok) {
  tok.clear();
  int c;
  while ((c = in.get()) != EOF) {
    if (c == '#') {
      while ((c = in.get()) != EOF && c != '\n') {}
      continue;
    }
    if (!std::isspace(c)) break;
  }
  if (c == EOF) return false;
  do tok.push_back((char)c); while ((c = in.get()) != EOF && !std::isspace(c));
  return true;   // the single whitespace after the last header token is consumed
}

pm::Mat read_pgm(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  std::string magic, w, h, maxval;
  if (!in || !pgm_token(in, magic) || magic != "P5" || !pgm_token(in, w) || !pgm_token(in, h) ||
      !pgm_token(in, maxval) || std::stoi(maxval) != 255)
    throw std::runtime_error(path + ": not an 8-bit binary PGM");
  pm::Mat m(std::stoi(h), std::stoi(w), pm::CV_8U);
  in.read(reinterpret_cast<char*>(m.data), (std::streamsize)m.total());
  if (!in) throw std::runtime_error(path + ": truncated");
  return m;
}

void write_npy(const std::string& path, const pm::Mat& m) {
  std::string hdr = "{'descr': '<f4', 'fortran_order': False, 'shape': (" +
                    std::to_string(m.rows) + ", " + std::to_string(m.cols) + "), }";
  const size_t pre = 10;   // magic (6) + version (2) + header length (2)
  hdr.append(64 - (pre + hdr.size() + 1) % 64, ' ');
  hdr.push_back('\n');
  std::ofstream out(path, std::ios::binary);
  out.write("\x93NUMPY\x01\x00", 8);
  const uint16_t n = (uint16_t)hdr.size();
  const char len[2] = {(char)(n & 0xFF), (char)(n >> 8)};
  out.write(len, 2);
  out.write(hdr.data(), (std::streamsize)hdr.size());
  for (int i = 0; i < m.rows; ++i)
    out.write(reinterpret_cast<const char*>(m.data + i * m.step), (std::streamsize)m.cols * 4);
  if (!out) throw std::runtime_error(path + ": write failed");
}

struct Message {
  int64_t t;
  int kind;   // 0 = pose, 1 = image (poses first at equal stamps)
  int index;
};

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> pos;
  pm::DriverConfig cfg;
  bool plan_only = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error(a + " needs a value");
      return argv[++i];
    };
    try {
      if (a == "--fx") cfg.fx = std::stod(val());
      else if (a == "--fy") cfg.fy = std::stod(val());
      else if (a == "--cx") cfg.cx = std::stod(val());
      else if (a == "--cy") cfg.cy = std::stod(val());
      else if (a == "--agg-num") cfg.agg_num = std::stoi(val());
      else if (a == "--interval") cfg.interval = std::stoi(val());
      else if (a == "--planes") cfg.planes = std::stoi(val());
      else if (a == "--dep-sample") cfg.dep_sample = std::stof(val());
      else if (a == "--plan-only") plan_only = true;
      else if (a == "--no-frame-ring") cfg.frame_ring = false;
      else pos.push_back(a);
    } catch (const std::exception& e) {
      std::cerr << "pm_evaluate: " << e.what() << "\n";
      return 2;
    }
  }
  if (pos.size() < 3 || pos.size() > 4) {
    std::cerr << "usage: pm_evaluate IMAGES POSES OUT_DIR [CASE] [--fx F] [--fy F] [--cx F] "
                 "[--cy F] [--agg-num N] [--interval N] [--planes N] [--dep-sample F] "
                 "[--plan-only] [--no-frame-ring]\n";
    return 2;
  }
  if (pos.size() == 4) CASE = std::atoi(pos[3].c_str());
  try {
    const std::string dir = pos[0].find('/') == std::string::npos
                                ? std::string(".")
                                : pos[0].substr(0, pos[0].rfind('/'));
    std::vector<std::pair<int64_t, std::string>> images;
    std::vector<std::pair<int64_t, std::vector<double>>> poses;
    std::string line;
    std::ifstream fi(pos[0]);
    if (!fi) throw std::runtime_error(pos[0] + ": cannot open");
    while (std::getline(fi, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ss(line);
      std::string st, path;
      int64_t t;
      if (!(ss >> st >> path) || !parse_stamp(st, &t))
        throw std::runtime_error(pos[0] + ": bad line: " + line);
      images.push_back({t, path[0] == '/' ? path : dir + "/" + path});
    }
    std::ifstream fp(pos[1]);
    if (!fp) throw std::runtime_error(pos[1] + ": cannot open");
    while (std::getline(fp, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ss(line);
      std::string st;
      std::vector<double> p(7);
      int64_t t;
      if (!(ss >> st) || !parse_stamp(st, &t))
        throw std::runtime_error(pos[1] + ": bad line: " + line);
      for (double& v : p)
        if (!(ss >> v)) throw std::runtime_error(pos[1] + ": bad line: " + line);
      poses.push_back({t, p});
    }
    std::vector<Message> msgs;
    for (size_t i = 0; i < poses.size(); ++i) msgs.push_back({poses[i].first, 0, (int)i});
    for (size_t i = 0; i < images.size(); ++i) msgs.push_back({images[i].first, 1, (int)i});
    std::stable_sort(msgs.begin(), msgs.end(), [](const Message& a, const Message& b) {
      return a.t != b.t ? a.t < b.t : a.kind < b.kind;
    });

    pm::FrameDriver driver(cfg, !plan_only);
    std::ofstream list;
    if (!plan_only) {
      list.open(pos[2] + "/depths.txt");
      if (!list) throw std::runtime_error(pos[2] + "/depths.txt: cannot create");
    }
    int produced = 0;
    for (const Message& m : msgs) {
      if (m.kind == 0) {
        const std::vector<double>& p = poses[m.index].second;
        driver.grab_pose(m.t, p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
      } else {
        driver.grab_image(m.t, read_pgm(images[m.index].second));
      }
      pm::DriverJob job;
      pm::Mat depth;
      if (!driver.process(job, plan_only ? nullptr : &depth)) continue;
      std::ostringstream desc;
      desc << stamp_text(job.stamp_ns);
      const std::string file = stamp_text(job.stamp_ns) + ".npy";
      if (!plan_only) {
        write_npy(pos[2] + "/" + file, depth);
        desc << " " << file;
      }
      desc << " " << job.ref_image;
      for (int s : job.src_images) desc << " " << s;
      if (plan_only) std::cout << desc.str() << "\n";
      else list << desc.str() << "\n";
      ++produced;
    }
    std::cerr << "pm_evaluate: " << produced << " depth maps, " << driver.skipped()
              << " steps waiting for source poses\n";
  } catch (const std::exception& e) {
    std::cerr << "pm_evaluate: " << e.what() << "\n";
    return 1;
  }
  return 0;
}

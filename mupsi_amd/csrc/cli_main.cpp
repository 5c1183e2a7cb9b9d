// mupsi — headless CLI over the GPU renderer, with the reference's contract
// (src/main.cpp:11-48): `mupsi [scene.json] [out.png]`, defaults scene.json and
// test.png; the JSON "output" key is ignored in favour of argv[2].
// Output as Framebuffer::save (src/rendering/framebuffer.cpp:37-78): .png is
// Reinhard-tonemapped, clamped, truncated to 8 bits, rows flipped so that camera
// row 0 is the bottom of the image; .pfm/.hdr keep the float radiance.
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mupsi.h"

namespace {

void put32(std::vector<unsigned char>& v, uint32_t x) {
  for (int s = 24; s >= 0; s -= 8) v.push_back((unsigned char)(x >> s));
}

void chunk(FILE* f, const char* type, const std::vector<unsigned char>& data) {
  std::vector<unsigned char> buf;
  put32(buf, (uint32_t)data.size());
  buf.insert(buf.end(), type, type + 4);
  buf.insert(buf.end(), data.begin(), data.end());
  const uint32_t crc = (uint32_t)crc32(0, buf.data() + 4, (uInt)(buf.size() - 4));
  put32(buf, crc);
  std::fwrite(buf.data(), 1, buf.size(), f);
}

bool write_png(const std::string& path, int w, int h, const std::vector<unsigned char>& rgb) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  std::fwrite(sig, 1, 8, f);
  std::vector<unsigned char> ihdr;
  put32(ihdr, (uint32_t)w);
  put32(ihdr, (uint32_t)h);
  ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
  chunk(f, "IHDR", ihdr);
  std::vector<unsigned char> raw;
  raw.reserve((size_t)h * (3 * w + 1));
  for (int y = 0; y < h; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), rgb.begin() + (size_t)y * w * 3, rgb.begin() + (size_t)(y + 1) * w * 3);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) {
    std::fclose(f);
    return false;
  }
  z.resize(zlen);
  chunk(f, "IDAT", z);
  chunk(f, "IEND", {});
  return std::fclose(f) == 0;
}

bool write_pfm(const std::string& path, int w, int h, const std::vector<float>& img) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "PF\n%d %d\n-1.0\n", w, h);
  // PFM rows run bottom-up: camera row 0 first
  std::fwrite(img.data(), 4, img.size(), f);
  return std::fclose(f) == 0;
}

bool write_hdr(const std::string& path, int w, int h, const std::vector<float>& img) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
  for (int r = 0; r < h; ++r) {
    const int y = h - 1 - r;  // top row first
    for (int x = 0; x < w; ++x) {
      const float* c = &img[((size_t)y * w + x) * 3];
      const float m = std::max(c[0], std::max(c[1], c[2]));
      unsigned char e[4] = {0, 0, 0, 0};
      if (m > 1e-32f) {
        int ex;
        const float s = std::frexp(m, &ex) * 256.0f / m;
        e[0] = (unsigned char)(c[0] * s);
        e[1] = (unsigned char)(c[1] * s);
        e[2] = (unsigned char)(c[2] * s);
        e[3] = (unsigned char)(ex + 128);
      }
      std::fwrite(e, 1, 4, f);
    }
  }
  return std::fclose(f) == 0;
}

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::printf("mupsi (MI355X) — mu + psi\n");
  const std::string config = argc >= 2 ? argv[1] : "scene.json";
  const std::string out = argc >= 3 ? argv[2] : "test.png";
  mupsi_ctx* ctx = nullptr;
  if (mupsi_open(0, &ctx) != MUPSI_OK) {
    std::fprintf(stderr, "mupsi: %s\n", mupsi_last_error());
    return -1;
  }
  mupsi_scene* scene = nullptr;
  int32_t spp = 1;
  if (mupsi_scene_load_json(ctx, config.c_str(), &scene, &spp) != MUPSI_OK) {
    std::fprintf(stderr, "Failed to load %s: %s\n", config.c_str(), mupsi_last_error());
    mupsi_close(ctx);
    return -1;
  }
  int32_t w = 0, h = 0;
  mupsi_scene_size(scene, &w, &h);
  const auto t0 = std::chrono::steady_clock::now();
  const int chunk_passes = std::max(1, std::min<int>(spp, (int)((8 << 20) / ((int64_t)w * h))));
  for (int k = 0; k < spp; k += chunk_passes) {
    const int n = std::min(chunk_passes, spp - k);
    if (mupsi_render_passes(scene, k, n, 0, nullptr) != MUPSI_OK) {
      std::fprintf(stderr, "render failed: %s\n", mupsi_last_error());
      return -1;
    }
    std::fprintf(stderr, "progress: %d%%\r", (100 * (k + n)) / spp);
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("\nRender complete (%.3f s, %.2f Mpaths/s). Saving %s...\n", secs,
              (double)w * h * spp / secs / 1e6, out.c_str());
  std::vector<double> sums((size_t)w * h * 3);
  int64_t count = 0;
  mupsi_read_accum(scene, sums.data(), &count);
  std::vector<float> img(sums.size());
  for (size_t i = 0; i < sums.size(); ++i) img[i] = (float)(sums[i] / (double)std::max<int64_t>(1, count));
  bool ok;
  if (ends_with(out, ".pfm")) {
    ok = write_pfm(out, w, h, img);
  } else if (ends_with(out, ".hdr") || ends_with(out, ".exr")) {
    ok = write_hdr(out, w, h, img);
  } else {
    std::vector<unsigned char> rgb((size_t)w * h * 3);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x)
        for (int c = 0; c < 3; ++c) {
          const float v = img[((size_t)y * w + x) * 3 + c];
          const float t = std::clamp(v / (1.0f + v), 0.0f, 1.0f);
          rgb[((size_t)(h - 1 - y) * w + x) * 3 + c] = (unsigned char)(t * 255.0f);
        }
    ok = write_png(out, w, h, rgb);
  }
  mupsi_scene_free(scene);
  mupsi_close(ctx);
  if (!ok) {
    std::fprintf(stderr, "cannot write %s\n", out.c_str());
    return -1;
  }
  return 0;
}

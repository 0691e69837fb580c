// Host utilities of the PIV path: Parameter.dat and pair-index parsing,
// output naming, 8-bit TIFF strip I/O, Tecplot writer and the seeded
// synthetic particle-pair generator.  Plain C++ (no GPU calls).
#include "hostio.h"
#include "synth.h"
#include "../../include/piv_xcorr2d.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <stdexcept>
#include <vector>

namespace {

thread_local std::string g_io_error;

int io_fail(int code, const std::string& msg) {
  g_io_error = msg;
  return code;
}

// CStdioFile::ReadString: one line without its terminator; false at EOF when
// nothing was read.  '\r' before '\n' is dropped (text mode).
bool read_line(FILE* f, std::string& out) {
  out.clear();
  int ch;
  bool any = false;
  while ((ch = fgetc(f)) != EOF) {
    any = true;
    if (ch == '\n') break;
    out.push_back((char)ch);
  }
  if (!any) return false;
  if (!out.empty() && out.back() == '\r') out.pop_back();
  return true;
}

// _wtoi / _wtof: leading whitespace, then the longest numeric prefix; 0 if none.
int parse_int(const std::string& s) {
  const char* p = s.c_str();
  char* end = nullptr;
  errno = 0;
  long v = strtol(p, &end, 10);
  if (end == p) return 0;
  if (v > 2147483647L) v = 2147483647L;
  if (v < -2147483647L - 1) v = -2147483647L - 1;
  return (int)v;
}
float parse_float(const std::string& s) {
  const char* p = s.c_str();
  char* end = nullptr;
  double v = strtod(p, &end);
  if (end == p) return 0.f;
  return (float)v;
}

void copy_str(char* dst, size_t cap, const std::string& s) {
  size_t n = std::min(cap - 1, s.size());
  memcpy(dst, s.data(), n);
  dst[n] = 0;
}

// ------------------------------------------------------------------ TIFF
struct TiffInfo {
  int width = 0, height = 0, bps = 0, compression = 1;
  std::vector<uint32_t> offs, counts;
};

struct Reader {
  FILE* f;
  bool big;
  bool ok = true;
  uint32_t u16(long pos) {
    unsigned char b[2];
    if (fseek(f, pos, SEEK_SET) != 0 || fread(b, 1, 2, f) != 2) { ok = false; return 0; }
    return big ? (b[0] << 8 | b[1]) : (b[1] << 8 | b[0]);
  }
  uint32_t u32(long pos) {
    unsigned char b[4];
    if (fseek(f, pos, SEEK_SET) != 0 || fread(b, 1, 4, f) != 4) { ok = false; return 0; }
    return big ? ((uint32_t)b[0] << 24 | b[1] << 16 | b[2] << 8 | b[3])
               : ((uint32_t)b[3] << 24 | b[2] << 16 | b[1] << 8 | b[0]);
  }
  // value i of an IFD entry at `ent` (type 3 SHORT or 4 LONG)
  uint32_t value(long ent, int i) {
    uint32_t type = u16(ent + 2), count = u32(ent + 4);
    uint32_t sz = (type == 3) ? 2 : 4;
    long base = (count * sz <= 4) ? ent + 8 : (long)u32(ent + 8);
    return (type == 3) ? u16(base + (long)i * 2) : u32(base + (long)i * 4);
  }
};

// getImHeadTiff (ImageIOPkg.cpp:124-258): walks every IFD; the last IFD's
// fields win, as in the reference's loop.  Counts and offsets read from the
// file are checked against the file size before anything is sized by them,
// and a short read ends the walk.
int tiff_parse_checked(const char* path, TiffInfo& ti, FILE** keep);

int tiff_parse(const char* path, TiffInfo& ti, FILE** keep) {
  try {
    return tiff_parse_checked(path, ti, keep);
  } catch (const std::exception& e) {
    return io_fail(PIVX_ERR_FORMAT, std::string("malformed TIFF: ") + e.what());
  }
}

int tiff_parse_checked(const char* path, TiffInfo& ti, FILE** keep) {
  FILE* f = fopen(path, "rb");
  if (!f) return io_fail(PIVX_ERR_IO, std::string("cannot open ") + path);
  long fsize = -1;
  if (fseek(f, 0, SEEK_END) == 0) fsize = ftell(f);
  if (fsize < 8 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "short TIFF header"); }
  unsigned char hdr[4];
  if (fread(hdr, 1, 4, f) != 4) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "short TIFF header"); }
  bool big;
  if (hdr[0] == 'I' && hdr[1] == 'I') big = false;
  else if (hdr[0] == 'M' && hdr[1] == 'M') big = true;
  else { fclose(f); return io_fail(PIVX_ERR_FORMAT, "not a TIFF file"); }
  Reader r{f, big};
  if (r.u16(2) != 42) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "TIFF identifier is not 42"); }
  uint32_t ifd = r.u32(4);
  int guard = 0;
  bool bad = false;
  while (ifd != 0 && r.ok && !bad && guard++ < 1024) {
    if ((long long)ifd + 2 > fsize) { bad = true; break; }
    uint32_t n = r.u16(ifd);
    if ((long long)ifd + 2 + 12LL * n + 4 > fsize) { bad = true; break; }
    for (uint32_t i = 0; i < n && r.ok && !bad; ++i) {
      long ent = (long)ifd + 2 + 12 * (long)i;
      uint32_t tag = r.u16(ent), type = r.u16(ent + 2), count = r.u32(ent + 4);
      switch (tag) {
        case 256: ti.width = (int)r.value(ent, 0); break;
        case 257: ti.height = (int)r.value(ent, 0); break;
        case 258: ti.bps = (int)r.value(ent, 0); break;
        case 259: ti.compression = (int)r.value(ent, 0); break;
        case 273:
        case 279: {
          // a strip table of `count` SHORT/LONG entries must lie inside the file
          const long long sz = type == 3 ? 2 : 4;
          if (count == 0 || (long long)count * sz > fsize) { bad = true; break; }
          std::vector<uint32_t>& t = tag == 273 ? ti.offs : ti.counts;
          t.resize(count);
          for (uint32_t k = 0; k < count && r.ok; ++k) t[k] = r.value(ent, (int)k);
          break;
        }
        default: break;
      }
    }
    if (bad || !r.ok) break;
    ifd = r.u32((long)ifd + 2 + 12 * (long)n);
  }
  if (bad) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "TIFF directory points outside the file"); }
  if (!r.ok) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "truncated TIFF"); }
  if (keep) *keep = f;
  else fclose(f);
  return PIVX_OK;
}

// ------------------------------------------------------------------ RNG
struct Rng {  // xoshiro256** seeded by splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9e3779b97f4a7c15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double gauss() {  // Box-Muller
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

struct Flow : pivx::SynthFlow {
  void eval(double x, double y, double& u, double& v) const { pivx::synth_flow_eval(*this, x, y, u, v); }
};

Flow make_flow(const pivx_synth_params* sp) {
  Flow f{};
  f.kind = sp->flow;
  f.a0 = sp->a0;
  f.a1 = sp->a1;
  f.a2 = sp->a2;
  f.xc = (sp->nx - 1) * 0.5;
  f.yc = (sp->ny - 1) * 0.5;
  if (f.kind == 1) {
    // max over s of (1 - exp(-s^2))/s, s = r/rc, reached at s ~ 1.1209
    double best = 0;
    for (int i = 1; i < 40000; ++i) {
      double s = i * 1e-4;
      best = std::max(best, (1 - std::exp(-s * s)) / s);
    }
    f.gamma = 2 * M_PI * sp->a0 * sp->a1 / best;
  } else if (f.kind == 2) {
    Rng r(sp->seed ^ 0x5eedf10dULL);
    for (int m = 0; m < 8; ++m) {
      f.fp[m] = (double)((int)(r.uni() * 7) - 3) / sp->nx;
      f.fq[m] = (double)((int)(r.uni() * 7) - 3) / sp->ny;
      f.fa[m] = 2 * r.uni() - 1;
      f.fb[m] = 2 * r.uni() - 1;
      f.fphi[m] = 2 * M_PI * r.uni();
      f.fpsi[m] = 2 * M_PI * r.uni();
    }
    f.scale = 1;
    // largest displacement over an 8-px grid, rows split over host threads
    // (a maximum: the same value whatever the split)
    const int rows = (sp->ny + 7) / 8;
    const int nt = (int)std::max(1u, std::min({16u, std::thread::hardware_concurrency(), (unsigned)rows}));
    std::vector<double> part((size_t)nt, 1e-12);
    auto scan = [&](int t) {
      double m = 1e-12;
      for (int j = t; j < rows; j += nt)
        for (int x = 0; x < sp->nx; x += 8) {
          double u, v;
          f.eval(x, 8 * j, u, v);
          m = std::max(m, std::sqrt(u * u + v * v));
        }
      part[(size_t)t] = m;
    };
    std::vector<std::thread> workers;
    int spawned = 1;
    try {
      for (; spawned < nt; ++spawned) workers.emplace_back(scan, spawned);
    } catch (...) {  // no thread to be had: the rest of the rows here
    }
    for (int t = spawned; t < nt; ++t) scan(t);
    scan(0);
    for (auto& w : workers) w.join();
    double mx = 1e-12;
    for (double m : part) mx = std::max(mx, m);
    f.scale = sp->a0 / mx;
  }
  return f;
}

}  // namespace

namespace pivx {

void set_global_error(const char* msg) { g_io_error = msg; }
const char* global_error() { return g_io_error.c_str(); }

// "%.2f" of a float, byte for byte as printf / iostream fixed+setprecision(2)
// print it: rint(x * 100) in double is the correctly rounded (ties to even)
// value of x * 100 for every float x.  For x = k * 2^e (k < 2^24) the product
// is exact when e >= -1; otherwise x * 100 lies at least 2^e away from any
// tie while the double product's error is below 100 * k * 2^(e - 53) < 2^e.
// The sign is printed for every negative x, so values rounding to zero print
// as "-0.00" like printf.  Non-finite and huge values go through snprintf.
static inline char* fmt_fixed2(char* p, float xf) {
  const double x = xf;
  if (!(std::fabs(x) < 1e15)) return p + snprintf(p, 64, "%.2f", x);
  uint64_t v = (uint64_t)std::fabs(std::rint(x * 100.0));
  if (std::signbit(x)) *p++ = '-';
  uint64_t ip = v / 100;
  const int fr = (int)(v % 100);
  char tmp[24];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + ip % 10);
    ip /= 10;
  } while (ip);
  while (n) *p++ = tmp[--n];
  *p++ = '.';
  *p++ = (char)('0' + fr / 10);
  *p++ = (char)('0' + fr % 10);
  return p;
}

// write_vec_data_general's text body (cu:1865-1973): rows are formatted in
// parallel chunks (one per host thread) and written in order.
int write_tecplot(const char* path, int nx_vec, int ny_vec, const float* gx, const float* gy, const float* dx,
                  const float* dy, int K, bool raw, bool crlf) {
  FILE* f = fopen(path, "wb");
  if (!f) return io_fail(PIVX_ERR_IO, std::string("cannot open ") + path);
  const char* nl = crlf ? "\r\n" : "\n";
  const size_t nll = crlf ? 2 : 1;
  std::string head;
  char line[512];
  head += "TITLE = \"B00001\"";
  head += nl;
  if (raw) head += "VARIABLES = \"x\", \"y\", \"dx1\", \"dy1\", \"dx2\", \"dy2\", \"dx3\", \"dy3\", \"dx4\", \"dy4\"";
  else head += "VARIABLES = \"x\", \"y\", \"displacement-x\", \"displacement-y\"";
  head += nl;
  snprintf(line, sizeof(line), "ZONE T=\"Frame 0\", I=%d, J=%d%s", nx_vec, ny_vec, nl);
  head += line;
  const long long nt = (long long)nx_vec * ny_vec;
  const int ncol = raw ? 10 : 4;
  auto format_rows = [&](long long r0, long long r1, std::string& out) {
    out.resize((size_t)(r1 - r0) * (ncol * 24 + 4));
    char* p = &out[0];
    for (long long i = r0; i < r1; ++i) {
      p = fmt_fixed2(p, gx[i]);
      *p++ = '\t';
      p = fmt_fixed2(p, gy[i]);
      if (raw) {
        for (int k = 0; k < 4; ++k) {
          *p++ = '\t';
          p = fmt_fixed2(p, k < K ? dx[i * K + k] : 0.f);
          *p++ = '\t';
          p = fmt_fixed2(p, k < K ? dy[i * K + k] : 0.f);
        }
      } else {
        *p++ = '\t';
        p = fmt_fixed2(p, dx[i]);
        *p++ = '\t';
        p = fmt_fixed2(p, dy[i]);
      }
      memcpy(p, nl, nll);
      p += nll;
    }
    out.resize((size_t)(p - out.data()));
  };
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nchunk = (int)std::max<long long>(1, std::min<long long>(hw, nt / 4096));
  std::vector<std::string> parts(nchunk);
  std::vector<std::thread> workers;
  for (int c = 1; c < nchunk; ++c)
    workers.emplace_back(format_rows, nt * c / nchunk, nt * (c + 1) / nchunk, std::ref(parts[c]));
  format_rows(0, nt / nchunk, parts[0]);
  for (auto& w : workers) w.join();
  bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
  for (const auto& part : parts) ok = ok && fwrite(part.data(), 1, part.size(), f) == part.size();
  ok = (fclose(f) == 0) && ok;
  return ok ? PIVX_OK : io_fail(PIVX_ERR_IO, std::string("write failed: ") + path);
}

void synth_flow_of(const pivx_synth_params* sp, SynthFlow* out) { *out = make_flow(sp); }

}  // namespace pivx

extern "C" {

int pivx_parse_parameter_file(const char* path, pivx_parameter_file* o) {
  if (!path || !o) return io_fail(PIVX_ERR_PARAM, "null argument");
  FILE* f = fopen(path, "rb");
  if (!f) return io_fail(PIVX_ERR_IO, "Can't open parameter file!");
  std::vector<std::string> L;
  std::string s;
  while ((int)L.size() < 50 && read_line(f, s)) L.push_back(s);
  fclose(f);
  if (L.size() < 50) return io_fail(PIVX_ERR_FORMAT, "Parameter missing in Parameter.dat!");
  memset(o, 0, sizeof(*o));
  copy_str(o->prefix_f1, sizeof(o->prefix_f1), L[0]);
  copy_str(o->suffix_f1, sizeof(o->suffix_f1), L[1]);
  copy_str(o->prefix_f2, sizeof(o->prefix_f2), L[2]);
  copy_str(o->suffix_f2, sizeof(o->suffix_f2), L[3]);
  copy_str(o->dir_index, sizeof(o->dir_index), L[4]);
  copy_str(o->folder_result, sizeof(o->folder_result), L[5]);
  int k = 6;
  auto I = [&]() { return parse_int(L[k++]); };
  auto F = [&]() { return parse_float(L[k++]); };
  pivx_init_params& p = o->init;
  o->device_id = I();
  o->ifmhe = I();
  o->mhe_ws = I();
  o->mhe_thresh = F();
  p.nx = I();
  p.ny = I();
  p.ws_x1 = I();
  p.ws_y1 = I();
  p.spacing_x1 = I();
  p.spacing_y1 = I();
  p.npass1 = I();
  p.ws_x2 = I();
  p.ws_y2 = I();
  p.spacing_x2 = I();
  p.spacing_y2 = I();
  p.npass2 = I();
  p.nsec_cc = I();
  p.wx1_median = I();
  p.wy1_median = I();
  o->npass1_median = I();
  o->min_nneigh_p1 = I();
  o->residual_threshold_p1 = F();
  o->insert_threshold_p1 = F();
  p.wx1_smooth = I();
  p.wy1_smooth = I();
  o->npass1_smooth = I();
  p.wx2_median = I();
  p.wy2_median = I();
  o->npass2_median = I();
  o->min_nneigh_p2 = I();
  o->residual_threshold_p2 = F();
  o->insert_threshold_p2 = F();
  p.wx2_smooth = I();
  p.wy2_smooth = I();
  o->npass2_smooth = I();
  p.max_dispx_percent1 = F();
  p.max_dispy_percent1 = F();
  p.max_dispx_percent2 = F();
  p.max_dispy_percent2 = F();
  p.npeak_save = I();
  p.dispx_lim1 = F();
  p.dispx_lim2 = F();
  p.dispy_lim1 = F();
  p.dispy_lim2 = F();
  return PIVX_OK;
}

int pivx_read_pair_index(const char* path, int* idx, int max_n) {
  FILE* f = path ? fopen(path, "rb") : nullptr;
  if (!f) {
    io_fail(PIVX_ERR_IO, "Can't open image pair index file!");
    return -1;
  }
  std::string s;
  int n = 0;
  while (read_line(f, s)) {
    if (idx && n < max_n) idx[n] = parse_int(s);
    ++n;
  }
  fclose(f);
  return n;
}

int pivx_output_name(const pivx_parameter_file* pf, int pair, int which, char* out, size_t cap) {
  if (!pf || !out || cap == 0) return io_fail(PIVX_ERR_PARAM, "null argument");
  const pivx_init_params& p = pf->init;
  char name[1024];
  switch (which) {
    case 0:  // main.cpp:328
      snprintf(name, sizeof(name), "_vector_GPU_Grid1_%d-peak-candidate_3DCC-%d-%d-%dpass_4peaks_raw.dat", p.npeak_save,
               p.ws_x1, p.ws_y1, p.npass1);
      break;
    case 1:  // main.cpp:336-337
      snprintf(name, sizeof(name),
               "_vector_GPU_Grid1_%d-peak-candidate_3DCC-%d-%d-%dpass_median-%d-%d-%.1f-%dpass_smooth-%d-%d-%dpass.dat",
               p.npeak_save, p.ws_x1, p.ws_y1, p.npass1, p.wx1_median, p.wy1_median, (double)pf->insert_threshold_p1,
               pf->npass1_median, p.wx1_smooth, p.wy1_smooth, pf->npass1_smooth);
      break;
    case 2:  // main.cpp:346
      snprintf(name, sizeof(name), "_vector_GPU_Grid2_%d-peak-candidate_3DCC-%d-%d-%dpass_4peaks_raw.dat", p.npeak_save,
               p.ws_x2, p.ws_y2, p.npass2);
      break;
    case 3:  // main.cpp:354-355
      snprintf(name, sizeof(name),
               "_vector_GPU_Grid2_%d-peak-candidate_3DCC-%d-%d-%dpass_median-%d-%d-%.1f-%dpass_smooth-%d-%d-%dpass.dat",
               p.npeak_save, p.ws_x2, p.ws_y2, p.npass2, p.wx2_median, p.wy2_median, (double)pf->insert_threshold_p2,
               pf->npass2_median, p.wx2_smooth, p.wy2_smooth, pf->npass2_smooth);
      break;
    case 4:  // main.cpp:368-369
      snprintf(name, sizeof(name), "_vector_GPU_Grid2_%d-peak-candidate_3DCC-%d-%d-%dpass_median-%d-%d-%.1f-%dpass.dat",
               p.npeak_save, p.ws_x2, p.ws_y2, p.npass2, p.wx2_median, p.wy2_median, (double)pf->insert_threshold_p2,
               pf->npass2_median);
      break;
    default:
      return io_fail(PIVX_ERR_PARAM, "which must be 0..4");
  }
  int n = snprintf(out, cap, "%s%04d%s", pf->folder_result, pair, name);
  if (n < 0 || (size_t)n >= cap) return io_fail(PIVX_ERR_PARAM, "output name too long");
  return PIVX_OK;
}

int pivx_write_tecplot(const char* path, int nx_vec, int ny_vec, const float* gx, const float* gy, const float* dx,
                       const float* dy, int K, int raw, int crlf) {
  if (!path || !gx || !gy || !dx || !dy || nx_vec < 0 || ny_vec < 0 || (raw && K < 1))
    return io_fail(PIVX_ERR_PARAM, "bad arguments");
  return pivx::write_tecplot(path, nx_vec, ny_vec, gx, gy, dx, dy, raw ? K : 1, raw != 0, crlf != 0);
}

int pivx_tiff_info(const char* path, int* w, int* h, int* bps) {
  TiffInfo ti;
  int rc = tiff_parse(path, ti, nullptr);
  if (rc) return rc;
  if (w) *w = ti.width;
  if (h) *h = ti.height;
  if (bps) *bps = ti.bps;
  return PIVX_OK;
}

// ImReadTiff8 (ImageIOPkg.cpp:280-429): strips concatenated in order.
int pivx_tiff_read8(const char* path, uint8_t* buf, size_t cap) {
  TiffInfo ti;
  FILE* f = nullptr;
  int rc = tiff_parse(path, ti, &f);
  if (rc) return rc;
  if (ti.bps != 8) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "only 8-bit TIFF is read on this path"); }
  if (ti.compression != 1) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "compressed TIFF is not supported"); }
  if (ti.offs.size() != ti.counts.size() || ti.offs.empty()) { fclose(f); return io_fail(PIVX_ERR_FORMAT, "bad strip tables"); }
  size_t pos = 0;
  for (size_t i = 0; i < ti.offs.size(); ++i) {
    size_t n = std::min<size_t>(ti.counts[i], cap - std::min(cap, pos));
    if (fseek(f, (long)ti.offs[i], SEEK_SET) != 0 || fread(buf + pos, 1, n, f) != n) {
      fclose(f);
      return io_fail(PIVX_ERR_FORMAT, "truncated strip data");
    }
    pos += n;
  }
  fclose(f);
  // a frame shorter than the buffer would leave stale pixels behind
  if (pos < cap) return io_fail(PIVX_ERR_FORMAT, "TIFF strips hold fewer bytes than the frame");
  return PIVX_OK;
}

int pivx_tiff_write8(const char* path, const uint8_t* buf, int width, int height) {
  FILE* f = fopen(path, "wb");
  if (!f) return io_fail(PIVX_ERR_IO, std::string("cannot open ") + path);
  auto w16 = [&](uint32_t v) { unsigned char b[2] = {(unsigned char)v, (unsigned char)(v >> 8)}; fwrite(b, 1, 2, f); };
  auto w32 = [&](uint32_t v) {
    unsigned char b[4] = {(unsigned char)v, (unsigned char)(v >> 8), (unsigned char)(v >> 16), (unsigned char)(v >> 24)};
    fwrite(b, 1, 4, f);
  };
  const uint32_t nbytes = (uint32_t)width * (uint32_t)height;
  const int nent = 8;
  const uint32_t ifd = 8, data = ifd + 2 + 12 * nent + 4;
  fwrite("II", 1, 2, f);
  w16(42);
  w32(ifd);
  w16(nent);
  auto ent = [&](uint32_t tag, uint32_t type, uint32_t val) {
    w16(tag); w16(type); w32(1);
    if (type == 3) { w16(val); w16(0); } else w32(val);
  };
  ent(256, 4, (uint32_t)width);
  ent(257, 4, (uint32_t)height);
  ent(258, 3, 8);
  ent(259, 3, 1);
  ent(262, 3, 1);
  ent(273, 4, data);
  ent(278, 4, (uint32_t)height);
  ent(279, 4, nbytes);
  w32(0);
  bool ok = fwrite(buf, 1, nbytes, f) == nbytes;
  ok = (fclose(f) == 0) && ok;
  return ok ? PIVX_OK : io_fail(PIVX_ERR_IO, "write failed");
}

// Particle model: count = round(density*(nx+32)*(ny+32)) particles uniform
// over [-16, nx+16) x [-16, ny+16); each a point-sampled Gaussian
// A*exp(-((x-xp)^2 + (y-yp)^2) / (2 sigma^2)) evaluated at integer pixel
// coordinates within 4 sigma; background + N(0, noise) per pixel, rounded and
// clipped to [0, 255].  Frame 2 uses (xp, yp) + flow(xp, yp).
int pivx_synth_pair(const pivx_synth_params* sp, uint8_t* img1, uint8_t* img2) {
  if (!sp || !img1 || !img2 || sp->nx < 1 || sp->ny < 1) return io_fail(PIVX_ERR_PARAM, "bad synth arguments");
  const int nx = sp->nx, ny = sp->ny;
  Flow flow = make_flow(sp);
  Rng r(sp->seed);
  const long long np = llround((double)sp->density * (nx + 32.0) * (ny + 32.0));
  std::vector<float> px(np), py(np), pa(np), qx(np), qy(np);
  for (long long i = 0; i < np; ++i) {
    px[i] = (float)(-16.0 + r.uni() * (nx + 32.0));
    py[i] = (float)(-16.0 + r.uni() * (ny + 32.0));
    pa[i] = (float)(sp->imin + r.uni() * (sp->imax - sp->imin));
    double u, v;
    flow.eval(px[i], py[i], u, v);
    qx[i] = (float)(px[i] + u);
    qy[i] = (float)(py[i] + v);
  }
  std::vector<float> acc((size_t)nx * ny);
  const double s2 = 2.0 * (double)sp->sigma * sp->sigma;
  const int rad = (int)std::ceil(4.0 * sp->sigma);
  for (int frame = 0; frame < 2; ++frame) {
    uint8_t* out = frame == 0 ? img1 : img2;
    const float* X = frame == 0 ? px.data() : qx.data();
    const float* Y = frame == 0 ? py.data() : qy.data();
    for (size_t i = 0; i < acc.size(); ++i) acc[i] = (float)(sp->background + sp->noise_sigma * r.gauss());
    for (long long i = 0; i < np; ++i) {
      const int cx = (int)std::floor(X[i] + 0.5f), cy = (int)std::floor(Y[i] + 0.5f);
      for (int y = std::max(0, cy - rad); y <= std::min(ny - 1, cy + rad); ++y) {
        const double dy2 = (y - (double)Y[i]) * (y - (double)Y[i]);
        for (int x = std::max(0, cx - rad); x <= std::min(nx - 1, cx + rad); ++x) {
          const double dx = x - (double)X[i];
          acc[(size_t)y * nx + x] += (float)(pa[i] * std::exp(-(dx * dx + dy2) / s2));
        }
      }
    }
    for (size_t i = 0; i < acc.size(); ++i) {
      float v = std::nearbyint(acc[i]);
      out[i] = (uint8_t)std::min(255.f, std::max(0.f, v));
    }
  }
  return PIVX_OK;
}

int pivx_synth_flow(const pivx_synth_params* sp, const float* x, const float* y, int n, float* u, float* v) {
  if (!sp || (n > 0 && (!x || !y || !u || !v))) return io_fail(PIVX_ERR_PARAM, "bad arguments");
  Flow flow = make_flow(sp);
  for (int i = 0; i < n; ++i) {
    double a, b;
    flow.eval(x[i], y[i], a, b);
    u[i] = (float)a;
    v[i] = (float)b;
  }
  return PIVX_OK;
}

}  // extern "C"

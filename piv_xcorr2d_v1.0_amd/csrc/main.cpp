// piv_xcorr2d: command-line driver of the MI355X PIV path.
//
// Same contract as the reference's main (xcorr2d_V1.0/main.cpp:10-394):
// argv[1] names a 50-line Parameter.dat (default "Parameter.dat"); frames are
// <prefix><%04d pair><suffix> 8-bit TIFFs; per pair it runs the grid-1 passes,
// upsampling, the grid-2 passes, and writes the five Tecplot files named as in
// main.cpp:327-376.  Windows-only pieces (MessageBox, getchar, MFC strings)
// are replaced by messages on stderr and exit codes.
//
// When Parameter.dat line 8 (ifmhe) is 1, both frames of every pair are MHE
// enhanced on the GPU with the window and threshold of lines 9-10, as
// main.cpp:267-283 does.
//
// Pairs are pipelined: the next pair's TIFFs are read and the previous
// pair's files written on host threads while the GPU runs the current pair
// (captured once as a HIP graph and replayed).
//
// Extra flags (after the parameter file): --crlf writes CRLF line ends like
// the reference's Windows text mode; --device N overrides Parameter.dat line 7;
// --ignore-mhe runs with MHE switched off even when line 8 asks for it;
// --serial runs each pair's read, upload, compute and writes one after the
// other, as the reference does (for comparison).
#include "../../include/piv_xcorr2d.h"

#include <sys/stat.h>

#include <chrono>
#include <future>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static int die(const char* msg) {
  fprintf(stderr, "%s\n", msg);
  return 1;
}

static bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}
static bool folder_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

int main(int argc, char** argv) {
  std::string dir_para = "Parameter.dat";
  bool crlf = false, ignore_mhe = false, serial = false;
  int device_override = -1;
  int argi = 1;
  if (argi < argc && argv[argi][0] != '-') dir_para = argv[argi++];
  for (; argi < argc; ++argi) {
    if (!strcmp(argv[argi], "--crlf")) crlf = true;
    else if (!strcmp(argv[argi], "--ignore-mhe")) ignore_mhe = true;
    else if (!strcmp(argv[argi], "--serial")) serial = true;
    else if (!strcmp(argv[argi], "--device") && argi + 1 < argc) device_override = atoi(argv[++argi]);
    else {
      fprintf(stderr, "usage: %s [Parameter.dat] [--crlf] [--device N] [--ignore-mhe] [--serial]\n", argv[0]);
      return 2;
    }
  }

  pivx_parameter_file pf;
  int rc = pivx_parse_parameter_file(dir_para.c_str(), &pf);
  if (rc) return die(pivx_last_error(nullptr));  // "Can't open parameter file!" / "Parameter missing ..."
  const int device = device_override >= 0 ? device_override : pf.device_id;

  std::vector<int> pairidx;
  int n = pivx_read_pair_index(pf.dir_index, nullptr, 0);
  if (n < 0) return die("Can't open image pair index file!");
  if (n == 0) return die("No image pair index detected!");
  pairidx.resize(n);
  pivx_read_pair_index(pf.dir_index, pairidx.data(), n);

  if (!folder_exists(pf.folder_result) && mkdir(pf.folder_result, 0755) != 0)
    return die("Folder creation not successful!");

  char tmp[32];
  snprintf(tmp, sizeof(tmp), "%04d", pairidx[0]);
  std::string dir_img1 = std::string(pf.prefix_f1) + tmp + pf.suffix_f1;
  if (!file_exists(dir_img1)) return die("An image for crosss-correlation is missing!");
  int w = 0, h = 0, bps = 0;
  if (pivx_tiff_info(dir_img1.c_str(), &w, &h, &bps)) return die(pivx_last_error(nullptr));

  pivx_ctx* ctx = nullptr;
  if (pivx_init(&pf.init, device, &ctx) != PIVX_OK) {
    fprintf(stderr, "%s\n", pivx_last_error(nullptr));
    return die("Initialization of xcorr2d_multi class failed!");
  }
  if (pf.ifmhe == 1 && !ignore_mhe && pivx_set_mhe(ctx, 1, pf.mhe_ws, pf.mhe_thresh) != PIVX_OK) {
    fprintf(stderr, "%s\n", pivx_last_error(ctx));
    pivx_freemem(ctx);
    return die("Initialization of mhe_gpu class failed!");
  }
  const size_t npix = (size_t)pf.init.nx * pf.init.ny;
  pivx_filter_params f1{pf.npass1_median, pf.min_nneigh_p1, pf.residual_threshold_p1, pf.insert_threshold_p1,
                        pf.npass1_smooth};
  pivx_filter_params f2{pf.npass2_median, pf.min_nneigh_p2, pf.residual_threshold_p2, pf.insert_threshold_p2,
                        pf.npass2_smooth};
  const int nout = pf.init.npass2 > 0 ? 5 : 2;

  // Pipeline (the reference runs read -> upload -> compute -> write serially
  // per pair, main.cpp:248-376): while the GPU runs pair i, one host thread
  // reads the TIFFs of pair i+1 into the other page-locked frame slot and
  // another writes the five files of pair i-1 from its result snapshot.
  uint8_t* slot[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  pivx_results* res[2] = {nullptr, nullptr};
  bool ok = true;
  for (int s = 0; s < 2 && ok; ++s) {
    slot[s][0] = static_cast<uint8_t*>(pivx_host_alloc(npix));
    slot[s][1] = static_cast<uint8_t*>(pivx_host_alloc(npix));
    ok = slot[s][0] && slot[s][1] && pivx_results_create(ctx, &res[s]) == PIVX_OK;
  }
  if (!ok) fprintf(stderr, "host buffers could not be allocated: %s\n", pivx_last_error(ctx));
  auto frame_name = [&](int k, int which) {
    char t[32];
    snprintf(t, sizeof(t), "%04d", k);
    return which == 0 ? std::string(pf.prefix_f1) + t + pf.suffix_f1 : std::string(pf.prefix_f2) + t + pf.suffix_f2;
  };
  // read pair index k into frame slot s; returns an error text or ""
  auto read_pair = [&](int k, int s) -> std::string {
    const std::string a = frame_name(k, 0), b = frame_name(k, 1);
    if (!file_exists(a) || !file_exists(b)) return "An image for crosss-correlation is missing!";
    for (const std::string* fn : {&a, &b}) {
      int fw = 0, fh = 0, fb = 0;
      if (pivx_tiff_info(fn->c_str(), &fw, &fh, &fb)) return pivx_last_error(nullptr);
      if (fw != pf.init.nx || fh != pf.init.ny)
        return *fn + ": frame is " + std::to_string(fw) + "x" + std::to_string(fh) + ", Parameter.dat gives " +
               std::to_string(pf.init.nx) + "x" + std::to_string(pf.init.ny);
    }
    // the two frames are read concurrently
    auto second = std::async(std::launch::async, [&]() -> std::string {
      return pivx_tiff_read8(b.c_str(), slot[s][1], npix) ? pivx_last_error(nullptr) : "";
    });
    std::string e = pivx_tiff_read8(a.c_str(), slot[s][0], npix) ? pivx_last_error(nullptr) : "";
    const std::string e2 = second.get();
    return e.empty() ? e2 : e;
  };
  // write the files of pair index k from snapshot s
  auto write_pair = [&](int k, int s) -> std::string {
    const int sels[5] = {0, 1, 2, 3, 4};
    for (int w = 0; w < nout; ++w) {
      char path[4096];
      if (pivx_output_name(&pf, k, w, path, sizeof(path)) || pivx_results_write(res[s], path, sels[w], crlf ? 1 : 0))
        return std::string("xcor.write_vec_data_general failed! ") + pivx_last_error(nullptr);
    }
    return "";
  };
  std::future<std::string> reader, writer[2];
  std::vector<std::chrono::steady_clock::time_point> t_start(pairidx.size());
  auto finish_writer = [&](int s, size_t i) {
    if (!writer[s].valid()) return;
    const std::string e = writer[s].get();
    if (!e.empty()) {
      fprintf(stderr, "%s\n", e.c_str());
      ok = false;
    } else {
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start[i]).count();
      printf("Done! %g seconds\n", secs);
    }
  };
  std::vector<size_t> written_by(2, 0);
  if (ok && !serial) ok = pivx_set_graph_mode(ctx, 1) == PIVX_OK;
  if (ok && serial) {
    for (size_t i = 0; i < pairidx.size() && ok; ++i) {
      t_start[i] = std::chrono::steady_clock::now();
      printf("Processing pair %d: %zu out of %zu\n", pairidx[i], i + 1, pairidx.size());
      std::string e = read_pair(pairidx[i], 0);
      if (e.empty() && (pivx_prep_img_data(ctx, slot[0][0], slot[0][1], pf.init.nx) ||
                        pivx_run_pair(ctx, &f1, &f2) || pivx_results_fetch(ctx, res[0]) ||
                        pivx_results_wait(res[0])))
        e = pivx_last_error(ctx);
      if (e.empty()) e = write_pair(pairidx[i], 0);
      if (!e.empty()) {
        fprintf(stderr, "%s\n", e.c_str());
        ok = false;
      } else {
        printf("Done! %g seconds\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start[i]).count());
      }
    }
  }
  if (ok && !serial) reader = std::async(std::launch::async, read_pair, pairidx[0], 0);
  for (size_t i = 0; i < pairidx.size() && ok && !serial; ++i) {
    const int s = (int)(i & 1);
    t_start[i] = std::chrono::steady_clock::now();
    printf("Processing pair %d: %zu out of %zu\n", pairidx[i], i + 1, pairidx.size());
    const std::string re = reader.get();
    if (!re.empty()) {
      fprintf(stderr, "%s\n", re.c_str());
      ok = false;
      break;
    }
    finish_writer(s, written_by[s]);  // snapshot s is free again
    if (!ok) break;
    if (pivx_prep_img_data_async(ctx, slot[s][0], slot[s][1], pf.init.nx) || pivx_run_pair(ctx, &f1, &f2) ||
        pivx_results_fetch(ctx, res[s])) {
      fprintf(stderr, "%s\n", pivx_last_error(ctx));
      ok = false;
      break;
    }
    // the other frame slot was last read by pair i-1's upload, which finished
    // before its results (fetched in the previous iteration) landed
    if (pivx_results_wait(res[s ^ 1]) != PIVX_OK) {
      fprintf(stderr, "%s\n", pivx_last_error(nullptr));
      ok = false;
      break;
    }
    if (i + 1 < pairidx.size()) reader = std::async(std::launch::async, read_pair, pairidx[i + 1], s ^ 1);
    if (pivx_results_wait(res[s]) != PIVX_OK) {
      fprintf(stderr, "%s\n", pivx_last_error(nullptr));
      ok = false;
      break;
    }
    writer[s] = std::async(std::launch::async, write_pair, pairidx[i], s);
    written_by[s] = i;
  }
  for (int s = 0; s < 2; ++s) finish_writer(s, written_by[s]);
  if (reader.valid()) reader.wait();
  if (ok && !pairidx.empty()) {
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start[0]).count();
    printf("Total: %zu pairs in %g seconds (%g pairs/s)\n", pairidx.size(), secs, pairidx.size() / secs);
  }
  for (int s = 0; s < 2; ++s) {
    pivx_results_destroy(res[s]);
    pivx_host_free(slot[s][0]);
    pivx_host_free(slot[s][1]);
  }
  pivx_freemem(ctx);
  if (ok) {
    printf("NO ERROR!\n");
    return 0;
  }
  printf("ERROR!\n");
  return 1;
}

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
// Extra flags (after the parameter file): --crlf writes CRLF line ends like
// the reference's Windows text mode; --device N overrides Parameter.dat line 7;
// --ignore-mhe runs with MHE switched off even when line 8 asks for it.
#include "../../include/piv_xcorr2d.h"

#include <sys/stat.h>

#include <chrono>
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
  bool crlf = false, ignore_mhe = false;
  int device_override = -1;
  int argi = 1;
  if (argi < argc && argv[argi][0] != '-') dir_para = argv[argi++];
  for (; argi < argc; ++argi) {
    if (!strcmp(argv[argi], "--crlf")) crlf = true;
    else if (!strcmp(argv[argi], "--ignore-mhe")) ignore_mhe = true;
    else if (!strcmp(argv[argi], "--device") && argi + 1 < argc) device_override = atoi(argv[++argi]);
    else {
      fprintf(stderr, "usage: %s [Parameter.dat] [--crlf] [--device N] [--ignore-mhe]\n", argv[0]);
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
  std::vector<uint8_t> img1(npix), img2(npix);
  pivx_filter_params f1{pf.npass1_median, pf.min_nneigh_p1, pf.residual_threshold_p1, pf.insert_threshold_p1,
                        pf.npass1_smooth};
  pivx_filter_params f2{pf.npass2_median, pf.min_nneigh_p2, pf.residual_threshold_p2, pf.insert_threshold_p2,
                        pf.npass2_smooth};
  bool ok = true;
  for (size_t i = 0; i < pairidx.size() && ok; ++i) {
    auto t0 = std::chrono::steady_clock::now();
    printf("Processing pair %d: %zu out of %zu\n", pairidx[i], i + 1, pairidx.size());
    snprintf(tmp, sizeof(tmp), "%04d", pairidx[i]);
    std::string a = std::string(pf.prefix_f1) + tmp + pf.suffix_f1;
    std::string b = std::string(pf.prefix_f2) + tmp + pf.suffix_f2;
    if (!file_exists(a) || !file_exists(b)) {
      fprintf(stderr, "An image for crosss-correlation is missing!\n");
      ok = false;
      break;
    }
    if (pivx_tiff_read8(a.c_str(), img1.data(), npix) || pivx_tiff_read8(b.c_str(), img2.data(), npix)) {
      fprintf(stderr, "%s\n", pivx_last_error(nullptr));
      ok = false;
      break;
    }
    if (pivx_prep_img_data(ctx, img1.data(), img2.data(), pf.init.nx) || pivx_run_pair(ctx, &f1, &f2)) {
      fprintf(stderr, "%s\n", pivx_last_error(ctx));
      ok = false;
      break;
    }
    const int sels[5] = {0, 1, 2, 3, 4};
    const int nout = pf.init.npass2 > 0 ? 5 : 2;
    for (int k = 0; k < nout && ok; ++k) {
      char path[4096];
      if (pivx_output_name(&pf, pairidx[i], k, path, sizeof(path)) ||
          pivx_write_vec_data_general(ctx, path, sels[k], crlf ? 1 : 0)) {
        fprintf(stderr, "xcor.write_vec_data_general failed! %s\n", pivx_last_error(ctx));
        ok = false;
      }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (ok) printf("Done! %g seconds\n", secs);
  }
  pivx_freemem(ctx);
  if (ok) {
    printf("NO ERROR!\n");
    return 0;
  }
  printf("ERROR!\n");
  return 1;
}

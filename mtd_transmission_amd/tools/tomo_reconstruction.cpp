// tomo_reconstruction: drop-in for Tomo_prog/Reconstruction (installed as
// tomo_reconstruction, INSTALL.sh:24). Same argv and configuration files;
// reads parametres.raw, tab_posSpec.raw and UBornfinal_{Re,Im}CxCxA.raw from
// CHEMIN_RESULT and writes indice.tif / absorption.tif (float32 pages; in
// polar mode indice_cplx_hdf5.hdf instead),
// OTF3D.tif (EXPORT_OTF), centre_reconstruction.tif and slice.tiff. The
// compute runs on the GPU through mtd_ewald_accumulate and mtd_finalize.
#include <cmath>
#include <fstream>

#include "h5_volume.hpp"
#include "tomo_common.hpp"

using namespace tomo;

int main(int argc, char** argv) {
  const Args args = parse_args(argc, argv);
  if (args.help) {
    print_help();
    return 0;
  }
  // manip::manip (Reconstruction/src/manip.cpp:10-163)
  const std::string gui = gui_conf_path(args.conf);
  const std::string chemin_acquis = mtd::cfg_string(gui, "CHEMIN_ACQUIS");
  std::string chemin_result, rep_config;
  if (args.polar) {
    const std::string racine = mtd::cfg_string(gui, "CHEMIN_RACINE");
    rep_config = racine + "/demosaic/" + args.etat + "UBorn/";
    chemin_result = rep_config;
  } else {
    chemin_result = mtd::cfg_string(gui, "CHEMIN_RESULT");
    rep_config = mtd::cfg_string(gui, "CHEMIN_CONFIG");
  }
  const std::string recon = rep_config + "/recon.txt";
  const std::string cfg = rep_config + "/config_manip.txt";

  double param[5];
  const std::string ppath = chemin_result + "/parametres.raw";
  if (mtd::file_size(ppath) < 40 || !mtd::read_raw(ppath, param, sizeof param)) {
    std::cerr << "Impossible d'ouvrir le fichier\n" << ppath << "\n";
    return 1;
  }
  const double nxmax = (double)(size_t)param[0];
  const int dim_roi = (int)param[3];
  auto val = [](const std::string& f, const char* k) { return (double)mtd::cfg_val(f, k, false, 0); };
  const double n0 = val(cfg, "N0"), na = val(cfg, "NA");
  const double coef = mtd::cfg_val(recon, "COEF_NA_OBJ_LIMIT", true, 1.0);
  const double lambda = val(cfg, "LAMBDA"), f_tube = val(cfg, "F_TUBE"), f_obj = val(cfg, "F_OBJ");
  const double tpcam = val(cfg, "TPCAM"), rf = val(cfg, "RF");
  const double dim_final = val(recon, "DIM_FINAL");
  const int born = val(recon, "BORN") != 0, c_aber = val(recon, "C_ABER") != 0;
  const bool export_otf = val(recon, "EXPORT_OTF") != 0;
  if ((size_t)nxmax == 0)
    std::cout << "----- ERREUR : NXMAX absent du fichier de configuration" << cfg << "\n";
  mtd_geom g;
  if (mtd_geom_derive(dim_roi, nxmax, dim_final, 0, 0, na, n0, lambda, tpcam, f_tube, f_obj, rf,
                      coef, born, c_aber, 0, &g) != MTD_OK) {
    std::cerr << "invalid geometry\n";
    return 1;
  }
  const int C = 2 * g.nxmax, V = g.dim_final;
  if (V < 4 * g.nxmax) std::cout << "Attention, dim_final<4*NXMAX\n";
  const double pix_tomo = g.pix_uborn * (2 * g.nxmax) / V;
  {
    std::ofstream f(chemin_result + "/SAV_param_manip.txt");
    f << "Fichiers d'acquisition : " << chemin_acquis << "\n";
    f << "Tp Holo " << g.pix_holo * 1e9 << " nm\nrayon " << g.rayon << " pixels\n";
    f << "Chp tomo pixel " << V << " pix^3\nDelta_fUborn " << g.delta_f << " m-1\n";
    f << "Tp Tomo " << pix_tomo * 1e9 << " nm\n";
  }

  // angle count from tab_posSpec.raw (main.cpp:105-107)
  const std::string tpath = chemin_result + "/tab_posSpec.raw";
  const long long tbytes = mtd::file_size(tpath);
  if (tbytes < 0) {
    std::cerr << "Impossible d'ouvrir le fichier\n" << tpath << "\n";
    return 1;
  }
  const int A = (int)(tbytes / 16);
  std::vector<double> tab(2 * (size_t)A);
  if (A > 0 && !mtd::read_raw(tpath, tab.data(), tab.size() * 8)) {
    std::cerr << "Problème lors de la lecture du fichier " << tpath << "\n";
    return 1;
  }
  std::vector<int32_t> pos(2 * (size_t)A);
  std::vector<float> centres((size_t)C * C, 0.f);
  for (int a = 0; a < A; ++a) {
    pos[a] = (int)tab[a];
    pos[a + A] = (int)tab[a + A];
    const long cpt = std::lround(tab[a + A]) * C + std::lround(tab[a]);
    if (cpt >= 0 && cpt < (long)centres.size()) centres[cpt] = (float)a;
  }
  // the reference joins this path without a separator (main.cpp:128)
  mtd::write_tiff_f32(chemin_result + "centre_reconstruction.tif", centres.data(), C, C, 1, 1.0, 1);

  const std::string dimImg = std::to_string(C) + "x" + std::to_string(C) + "x" + std::to_string(A);
  const size_t np = (size_t)A * C * C;
  std::vector<double> re(np), im(np);
  const std::string rpath = chemin_result + "/UBornfinal_Re" + dimImg + ".raw";
  const std::string ipath = chemin_result + "/UBornfinal_Im" + dimImg + ".raw";
  if (mtd::file_size(rpath) != (long long)(np * 8) || mtd::file_size(ipath) != (long long)(np * 8)) {
    std::cerr << "Taille du fichier " << rpath << " incompatible avec les dimensions\n";
    return 1;
  }
  mtd::read_raw(rpath, re.data(), np * 8);
  mtd::read_raw(ipath, im.data(), np * 8);
  std::vector<float> fields(2 * np);
  for (size_t i = 0; i < np; ++i) {
    fields[2 * i] = (float)re[i];
    fields[2 * i + 1] = (float)im[i];
  }
  std::vector<double>().swap(re);
  std::vector<double>().swap(im);

  mtd_ctx* ctx = nullptr;
  if (mtd_ctx_create(0, &ctx) != MTD_OK) {
    std::cerr << "no GPU available\n";
    return 1;
  }
  const size_t nv = (size_t)V * V * V;
  void *d_f = nullptr, *d_p = nullptr, *d_vol = nullptr, *d_cnt = nullptr, *d_re = nullptr,
       *d_im = nullptr, *d_otf = nullptr;
  hip_or_die(hipMalloc(&d_f, fields.size() * 4 + 8), "hipMalloc");
  hip_or_die(hipMalloc(&d_p, pos.size() * 4 + 8), "hipMalloc");
  hip_or_die(hipMalloc(&d_vol, nv * 8), "hipMalloc");
  hip_or_die(hipMalloc(&d_cnt, nv * 4), "hipMalloc");
  hip_or_die(hipMalloc(&d_re, nv * 4), "hipMalloc");
  hip_or_die(hipMalloc(&d_im, nv * 4), "hipMalloc");
  if (export_otf) hip_or_die(hipMalloc(&d_otf, nv), "hipMalloc");
  hip_or_die(hipMemcpy(d_f, fields.data(), fields.size() * 4, hipMemcpyHostToDevice), "upload");
  hip_or_die(hipMemcpy(d_p, pos.data(), pos.size() * 4, hipMemcpyHostToDevice), "upload");
  Clock t1;
  mtd_or_die(ctx, mtd_ewald_accumulate(ctx, &g, (const float*)d_f, (const int32_t*)d_p, A,
                                       (float*)d_vol, (int32_t*)d_cnt),
             "Ewald accumulation");
  mtd_or_die(ctx, mtd_ctx_sync(ctx), "sync");
  std::cout << "Temps total fft+retroprop= " << t1.s() << "\n";
  Clock t2;
  mtd_or_die(ctx, mtd_finalize(ctx, &g, (const float*)d_vol, (const int32_t*)d_cnt, (float*)d_re,
                               (float*)d_im, (uint8_t*)d_otf),
             "finalize");
  mtd_or_die(ctx, mtd_ctx_sync(ctx), "sync");
  std::cout << "temps TF3D= " << t2.s() * 1e3 << " ms\n";
  std::vector<float> ire(nv), iim(nv);
  hip_or_die(hipMemcpy(ire.data(), d_re, nv * 4, hipMemcpyDeviceToHost), "download");
  hip_or_die(hipMemcpy(iim.data(), d_im, nv * 4, hipMemcpyDeviceToHost), "download");
  std::vector<float> otf;
  if (export_otf) {
    std::vector<uint8_t> o(nv);
    hip_or_die(hipMemcpy(o.data(), d_otf, nv, hipMemcpyDeviceToHost), "download");
    otf.assign(o.begin(), o.end());
  }
  for (void* p : {d_f, d_p, d_vol, d_cnt, d_re, d_im, d_otf}) (void)hipFree(p);
  mtd_ctx_destroy(ctx);

  Clock t3;
  bool ok = true;
  if (export_otf)
    ok &= mtd::write_tiff_f32(chemin_result + "/OTF3D.tif", otf.data(), V, V, V, pix_tomo, 4);
  // polar mode: HDF5 instead of the two TIFFs (main.cpp:360-367)
  bool h5_done = false;
  if (args.polar) {
    const std::string h5path = chemin_result + "/indice_cplx_hdf5.hdf";
    std::cout << "sauvegarde en hdf dans" << h5path << "\n";
    h5_done = write_complex_volume_h5(h5path, ire.data(), iim.data(), V, V, V);
    if (h5_done)
      std::cout << "sauvegarde HDF!\n";
    else
      std::cout << "mode polar: libhdf5 not loadable, volume written as TIFF instead\n";
  }
  if (!h5_done) {
    ok &= mtd::write_tiff_f32(chemin_result + "/indice.tif", ire.data(), V, V, V, pix_tomo, 4);
    ok &= mtd::write_tiff_f32(chemin_result + "/absorption.tif", iim.data(), V, V, V, pix_tomo, 4);
  }
  // extractSliceZ(indice, "x", 236) (main.cpp:375-376, fonctions.cpp:18-52)
  const int sx = 236;
  if (sx < V) {
    std::vector<float> slice((size_t)V * V);
    for (int z = 0; z < V; ++z)
      for (int y = 0; y < V; ++y) slice[(size_t)y + (size_t)z * V] = ire[sx + (size_t)y * V + (size_t)z * V * V];
    ok &= mtd::write_tiff_f32(chemin_result + "/slice.tiff", slice.data(), V, V, 1, pix_tomo, 1);
  } else {
    std::cout << "slice.tiff skipped: x=236 outside a " << V << "^3 volume\n";
  }
  std::cout << "Temps total ecriture tiff= " << t3.s() << "\n";
  if (!ok) {
    std::cerr << "Erreur d'écriture dans " << chemin_result << "\n";
    return 1;
  }
  std::cout << "#######-- Reconstruction terminée --#######\nRésultats dans " << chemin_result
            << "\n";
  return 0;
}

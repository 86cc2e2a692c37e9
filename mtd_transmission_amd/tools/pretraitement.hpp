// The preprocessing program shared by the two installed names,
// tomo_pretraitement (pretraitement_hors_axe_CPU) and tomo_pretraitement_GPU
// (pretraitement_hors_axe_GPU): same configuration files ($HOME/.config/
// gui_tomo[_polar].conf -> recon.txt, config_manip.txt), same outputs in
// CHEMIN_RESULT: UBornfinal_{Re,Im}CxCxA.raw (f64), parametres.raw,
// tab_posSpec.raw, centres.tif, SAV_param_manip.txt; the compute runs on the
// GPU through mtd_preprocess. The variants differ in their options (Pre).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>

#include "tomo_common.hpp"

namespace tomo {

// What sets the two programs apart.
struct Pre {
  std::string conf = "gui_tomo.conf";  // configuration name under $HOME/.config
  bool polar = false;
  std::string etat;
  // the ArrayFire variant's windows (pretraitement_hors_axe_GPU/main.cpp:
  // 55-60): tukey alpha 0.025 for the hologram and a (1+i) tukey2D(C, 0.025)
  // crop window
  bool arrayfire = false;
  // the ArrayFire variant's extra outputs: phasePI_atan2.raw, the wrapped
  // phase atan2(U) of every angle's demodulated field (float32, appended,
  // main.cpp:212-214, only with C_ABER 1), and UBornAmpFinal.raw, the
  // corrected amplitude of the last angle (float32, appended, main.cpp:319)
  bool variant_outputs = false;
  int device = 0;  // GPU of the context
};

inline int run_pretraitement(const Pre& args) {
  // manip::manip (pretraitement_hors_axe_CPU/manip.cpp:9-157)
  const std::string gui = gui_conf_path(args.conf);
  const std::string rep_config = mtd::cfg_string(gui, "CHEMIN_CONFIG");
  std::string chemin_result = mtd::cfg_string(gui, "CHEMIN_RESULT") + "/";
  std::string chemin_acquis = mtd::cfg_string(gui, "CHEMIN_ACQUIS") + "/";
  const std::string chemin_racine = mtd::cfg_string(gui, "CHEMIN_RACINE") + "/";
  const std::string recon = rep_config + "/recon.txt";
  const std::string cfg = rep_config + "/config_manip.txt";
  auto val = [](const std::string& f, const char* k) { return (double)mtd::cfg_val(f, k, false, 0); };
  const double n0 = val(cfg, "N0"), na = val(cfg, "NA"), lambda = val(cfg, "LAMBDA");
  const double f_tube = val(cfg, "F_TUBE"), f_obj = val(cfg, "F_OBJ");
  const double tpcam = val(cfg, "TPCAM"), rf = val(cfg, "RF");
  const double nxmax = val(cfg, "NXMAX"), ccd = val(recon, "CCD_ROIX");
  const double dim_final = val(recon, "DIM_FINAL");
  const double cx = val(cfg, "CIRCLE_CX"), cy = val(cfg, "CIRCLE_CY");
  const long premier = (long)(size_t)val(recon, "PREMIER_ANGLE");
  const long nb_holo = (long)(size_t)val(cfg, "NB_HOLO");
  const int c_aber = val(recon, "C_ABER") != 0;
  // DEROUL / VOLKOV are read as booleans (manip.cpp:77-79); VOLKOV 0 selects
  // Herraez
  const int deroul = val(recon, "DEROUL") == 0 ? MTD_DEROUL_OFF
                     : val(recon, "VOLKOV") != 0 ? MTD_DEROUL_VOLKOV : MTD_DEROUL_HERRAEZ;
  const int born = val(recon, "BORN") != 0;
  if ((size_t)nxmax == 0)
    std::cout << "----- ERREUR : NXMAX absent du fichier de configuration" << cfg << "\n";

  std::string mask_path;
  if (args.polar) {
    chemin_acquis = chemin_racine + "demosaic" + args.etat;
    chemin_result = chemin_racine + "demosaic" + args.etat + "/UBorn/";
    mask_path = chemin_racine + "Image_mask.pgm";
  } else {
    mask_path = chemin_acquis + "/Image_mask.pgm";
  }
  mtd_geom g;
  if (mtd_geom_derive(ccd, nxmax, dim_final, cx, cy, na, n0, lambda, tpcam, f_tube, f_obj, rf,
                      1.0, born, c_aber, deroul, &g) != MTD_OK) {
    std::cerr << "invalid geometry (NXMAX)\n";
    return 1;
  }
  if (args.arrayfire) {  // pretraitement_hors_axe_GPU/main.cpp:55-60
    g.alpha_holo = 0.025;
    g.alpha_crop = 0.025;
  }
  const int N = g.n, C = 2 * g.nxmax;
  const std::string sav_param = chemin_result + "/SAV_param_manip.txt";
  {
    std::ofstream f(sav_param);
    f << "Fichiers d'acquisition : " << chemin_acquis << "\n";
    f << "Tp Holo " << g.pix_holo * 1e9 << " nm\nChamp " << N * g.pix_holo * 1e6 << " um\n";
    f << "rayon " << g.rayon << " pixels\nNXMAX " << g.nxmax << "\nNA=" << na << "\n";
  }

  // hologram loop (main.cpp:130-165): missing files are skipped and the
  // angles compacted
  const long nb_angle = nb_holo - premier > 0 ? nb_holo - premier : 0;
  // pinned host stack: the library streams it to the GPU, overlapping the
  // copies with the hologram transforms
  uint16_t* holos = nullptr;
  if (nb_angle > 0)
    hip_or_die(hipHostMalloc((void**)&holos, (size_t)nb_angle * N * N * sizeof(uint16_t), 0),
               "hipHostMalloc");
  int A = 0;
  Clock t_load;
  for (long k = 0; k < nb_angle; ++k) {
    if (k % 100 == 0) std::cout << k << "\n";
    char num[16];
    snprintf(num, sizeof num, "%03ld", k);
    const std::string path = chemin_acquis + "/i" + num + ".pgm";
    if (!exists(path)) {
      std::cout << "fichier " << k << " inexistant\n";
      continue;
    }
    std::vector<uint16_t> px;
    int w, h, maxv;
    std::string err;
    if (!mtd::read_pgm(path, px, w, h, maxv, &err) || w < N || h < N) {
      std::cerr << "Impossible d'ouvrir " << path << " " << err << "\n";
      return 1;
    }
    for (int y = 0; y < N; ++y)
      std::copy(px.begin() + (size_t)y * w, px.begin() + (size_t)y * w + N,
                holos + ((size_t)A * N + y) * N);
    ++A;
  }
  std::cout << "NbAngle=" << nb_angle << " trouves=" << A << "\n";
  std::cout << "Temps chargement hologrammes (disque)= " << t_load.s() << "\n";

  std::vector<uint8_t> mask;
  const uint8_t* maskp = nullptr;
  if (exists(mask_path)) {
    std::vector<uint16_t> px;
    int w, h, maxv;
    if (mtd::read_pgm(mask_path, px, w, h, maxv) && w == C && h == C) {
      mask.resize((size_t)C * C);
      for (size_t i = 0; i < mask.size(); ++i) mask[i] = (uint8_t)std::min<int>(px[i], 255);
      maskp = mask.data();
      std::ofstream(sav_param, std::ios::app)
          << "un masque d'aberration a été utilisé-------------------------------------------"
             "---------------------\n";
    } else {
      std::cout << "Problème aberrations : masque " << mask_path << " de largeur " << w
                << ", mais image de largeur " << C << "\n";
    }
  } else {
    std::cout << "Masque : " << mask_path << " non ---trouvé création masque unité\n";
  }
  std::ofstream(sav_param, std::ios::app) << "degré poly aberration=4\n";

  // GPU: holograms -> fields
  mtd_ctx* ctx = nullptr;
  if (mtd_ctx_create(args.device, &ctx) != MTD_OK) {
    std::cerr << "no GPU available\n";
    return 1;
  }
  std::vector<float> fields((size_t)A * C * C * 2);
  std::vector<int32_t> pos((size_t)2 * A);
  Clock t_gpu;
  if (A > 0) {
    void *d_f = nullptr, *d_p = nullptr;
    hip_or_die(hipMalloc(&d_f, fields.size() * 4), "hipMalloc");
    hip_or_die(hipMalloc(&d_p, pos.size() * 4), "hipMalloc");
    mtd_or_die(ctx, mtd_preprocess(ctx, &g, holos, MTD_U16, A, maskp, (int32_t*)d_p, (float*)d_f),
               "preprocess");
    mtd_or_die(ctx, mtd_ctx_sync(ctx), "sync");
    hip_or_die(hipMemcpy(fields.data(), d_f, fields.size() * 4, hipMemcpyDeviceToHost), "download");
    hip_or_die(hipMemcpy(pos.data(), d_p, pos.size() * 4, hipMemcpyDeviceToHost), "download");
    if (args.variant_outputs && c_aber) {
      // the wrapped phase of each demodulated field, from the crops and
      // speculars of the same run
      std::vector<float> u(fields.size());
      void* d_c = nullptr;
      hip_or_die(hipMalloc(&d_c, u.size() * 4), "hipMalloc");
      mtd_or_die(ctx, mtd_holo_to_crop(ctx, &g, holos, MTD_U16, A, (float*)d_c), "crops");
      mtd_or_die(ctx, mtd_demodulate(ctx, &g, (const float*)d_c, A, (const int32_t*)d_p,
                                     (float*)d_f),
                 "demodulation");
      mtd_or_die(ctx, mtd_ctx_sync(ctx), "sync");
      hip_or_die(hipMemcpy(u.data(), d_f, u.size() * 4, hipMemcpyDeviceToHost), "download");
      (void)hipFree(d_c);
      std::vector<float> phase((size_t)A * C * C);
      for (size_t i = 0; i < phase.size(); ++i) phase[i] = std::atan2(u[2 * i + 1], u[2 * i]);
      if (!mtd::write_raw(chemin_result + "/phasePI_atan2.raw", phase.data(), phase.size() * 4,
                          true)) {
        std::cerr << "Erreur d'écriture de phasePI_atan2.raw\n";
        return 1;
      }
    }
    (void)hipFree(d_f);
    (void)hipFree(d_p);
  }
  if (holos) (void)hipHostFree(holos);
  std::cout << "Temps pour FFT holo+découpe Spectre+part2 (GPU)= " << t_gpu.s() << "\n";

  // outputs (main.cpp:171-173,325-326,351-358)
  Clock t_write;
  const std::string dimImg = std::to_string(C) + "x" + std::to_string(C) + "x" + std::to_string(A);
  const size_t np = (size_t)A * C * C;
  std::vector<double> re(np), im(np);
  for (size_t i = 0; i < np; ++i) {
    re[i] = fields[2 * i];
    im[i] = fields[2 * i + 1];
  }
  const bool ok1 = mtd::write_raw(chemin_result + "/UBornfinal_Re" + dimImg + ".raw", re.data(),
                                  np * 8, false);
  const bool ok2 = mtd::write_raw(chemin_result + "/UBornfinal_Im" + dimImg + ".raw", im.data(),
                                  np * 8, false);
  std::vector<double> param{(double)g.nxmax, (double)A, g.rayon, (double)N, g.pix_holo};
  const bool ok3 = mtd::write_raw(chemin_result + "/parametres.raw", param.data(), 5 * 8, false);
  std::vector<double> tab(2 * (size_t)A);
  std::vector<float> centre((size_t)C * C, 0.f);
  for (int a = 0; a < A; ++a) {
    tab[a] = pos[a];
    tab[a + A] = pos[a + A];
    centre[(size_t)pos[a] * C + pos[a + A]] = (float)a;  // main.cpp:252 (x-major index)
  }
  const bool ok4 = mtd::write_raw(chemin_result + "/tab_posSpec.raw", tab.data(), tab.size() * 8,
                                  false);
  const bool ok5 = mtd::write_tiff_f32(chemin_result + "/centres.tif", centre.data(), C, C, 1,
                                       g.na / g.nxmax, 1);
  bool ok6 = true;
  if (args.variant_outputs) {
    // the corrected amplitude |A| of the last angle (main.cpp:274-279,319):
    // Born fields are |A| e^(i phi) - 1, Rytov fields log|A| + i phi; zeros
    // without the aberration correction, as the variant's unused buffer
    std::vector<float> amp((size_t)C * C, 0.f);
    if (A > 0 && c_aber) {
      const float* f = fields.data() + (size_t)(A - 1) * C * C * 2;
      for (size_t i = 0; i < amp.size(); ++i)
        amp[i] = born ? std::hypot(f[2 * i] + 1.0f, f[2 * i + 1]) : std::exp(f[2 * i]);
    }
    ok6 = mtd::write_raw(chemin_result + "/UBornAmpFinal.raw", amp.data(), amp.size() * 4, true);
  }
  mtd_ctx_destroy(ctx);
  std::cout << "Temps ecriture (disque)= " << t_write.s() << "\n";
  if (!(ok1 && ok2 && ok3 && ok4 && ok5 && ok6)) {
    std::cerr << "Erreur d'écriture dans " << chemin_result << "\n";
    return 1;
  }
  std::cout << "Fin prétraitement\nResults in " << chemin_result << "\n";
  return 0;
}

}  // namespace tomo

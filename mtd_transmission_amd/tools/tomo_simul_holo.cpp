// tomo_simul_holo: drop-in for Tomo_prog/Simul_holo (its main.cpp), the
// synthetic-acquisition generator that feeds tomo_reconstruction. Same
// configuration ($HOME/.config/gui_tomo_polar.conf -> CHEMIN_CONFIG /
// config_manip.txt, CHEMIN_RESULT), same object (a rectangular block,
// genere_barre), same "flower" illumination pattern (OTF::bFleur, 4 axes) and
// the outputs tomo_reconstruction reads: UBornfinal_{Re,Im}CxCxA.raw,
// tab_posSpec.raw, parametres.raw, plus obj_Re.tif and centres_simul.tif. The
// spectra come from the library's Born forward model (mtd_forward_spectra:
// TF3Dcplx of the object, then calcHolo per angle) and the fields from
// mtd_spectra_to_fields, on the GPU.
//
// Not produced: the OTF/convolved-object diagnostics (OTF_simule_Re.tif,
// obj_conv_*.tif, TF_obj_Re.tif) and the Rytov phase raws, which no other
// program reads.
#include <cmath>
#include <fstream>

#include "tomo_common.hpp"

using namespace tomo;

// C round(): halves away from zero
static double round_away(double x) { return x < 0 ? -std::floor(-x + 0.5) : std::floor(x + 0.5); }

int main(int argc, char** argv) {
  const Args args = parse_args(argc, argv);
  if (args.help) {
    print_help();
    return 0;
  }
  // manip::manip (Simul_holo/src/manip.cpp:10-60); main.cpp:30 fixes the
  // hologram side at 2048
  const int dimROI = 2048;
  const std::string gui = gui_conf_path("gui_tomo_polar.conf");
  const std::string rep_config = mtd::cfg_string(gui, "CHEMIN_CONFIG");
  const std::string chemin_result = mtd::cfg_string(gui, "CHEMIN_RESULT");
  const std::string cfg = rep_config + "/config_manip.txt";
  auto val = [&](const char* k, bool has_def = false, float def = 0.f) {
    return (double)mtd::cfg_val(cfg, k, has_def, def);
  };
  const int nb_holo = (int)(uint16_t)val("NB_HOLO");
  const double n0 = val("N0"), na_obj = val("NA_OBJ", true, 1.3f), lambda = val("LAMBDA");
  const double f_tube = val("F_TUBE"), f_obj = val("F_OBJ"), rf = val("RF"), tpcam = val("TPCAM");
  const double gt = f_tube / (f_obj * rf);
  const double tp_holo = tpcam / gt;
  const double df_holo = 1.0 / (dimROI * tp_holo);
  const double theta_max = std::asin(na_obj / n0);
  const int nxmax = (int)(uint16_t)(na_obj / (lambda * df_holo));
  const int dim_final = 4 * nxmax, C = 2 * nxmax;
  const double rayon = round_away(nxmax / std::sin(theta_max));
  const double tp_uborn = tp_holo * dimROI / (2 * nxmax);
  const double tp_tomo = tp_uborn * (2 * nxmax) / dim_final;
  std::cout << "NXMAX=" << nxmax << " R_Ewald=" << rayon << " dim_final=" << dim_final
            << " Tp_Tomo=" << tp_tomo << "\n";
  if (nxmax <= 0 || nb_holo <= 0) {
    std::cerr << "invalid configuration (NXMAX or NB_HOLO)\n";
    return 1;
  }

  // the reconstruction's geometry of the same acquisition; the simulator's
  // R_EwaldPix = round(NXMAX / sin(asin(NA/n0)))
  mtd_geom g;
  if (mtd_geom_derive(dimROI, nxmax, dim_final, dimROI / 2, dimROI / 2, na_obj, n0, lambda, tpcam,
                      f_tube, f_obj, rf, 1.0, 1, 1, MTD_DEROUL_OFF, &g) != MTD_OK) {
    std::cerr << "invalid geometry\n";
    return 1;
  }
  g.rayon = rayon;

  // genere_barre (Simul_holo/src/fonctions.cpp:393-423) with main.cpp:57-66's
  // block: x, y in [-2.5, 2.5] um, z in [-2.5, 2.5] um, n = 1.395, kappa 0
  const int V = dim_final;
  const size_t nv = (size_t)V * V * V;
  std::vector<float> obj(nv, 0.f);
  const double dn = 1.395 - n0;
  const double x0 = -2.5e-6, x1 = x0 + 5e-6, y0 = -2.5e-6, y1 = y0 + 5e-6, zw = 5e-6;
  const int lo[3] = {(int)round_away(x0 / tp_tomo), (int)round_away(y0 / tp_tomo),
                     (int)round_away(-zw / 2 / tp_tomo)};
  const int hi[3] = {(int)round_away(x1 / tp_tomo), (int)round_away(y1 / tp_tomo),
                     (int)round_away(zw / 2 / tp_tomo)};
  for (int x = lo[0] + V / 2; x < hi[0] + V / 2; ++x)
    for (int y = lo[1] + V / 2; y < hi[1] + V / 2; ++y)
      for (int z = lo[2] + V / 2; z < hi[2] + V / 2; ++z)
        if (x >= 0 && x < V && y >= 0 && y < V && z >= 0 && z < V)
          obj[(size_t)z * V * V + (size_t)y * V + x] = (float)dn;

  // OTF::bFleur (Simul_holo/src/OTF.cpp:487-509), 4 axes: speculars in
  // centred coordinates, rounded; tab_posSpec in crop pixels (main.cpp:203-204)
  const int A = nb_holo;
  std::vector<int32_t> pos(2 * (size_t)A);
  std::vector<float> centres((size_t)C * C, 0.f);
  const double dtheta = 2 * M_PI / A;
  for (int k = 0; k < A; ++k) {
    const double th = dtheta * k;
    const double sx = round_away((nxmax - 1) * std::cos(4 * th) * std::cos(th));
    const double sy = round_away((nxmax - 1) * std::cos(4 * th) * std::sin(th));
    pos[k] = (int32_t)(sx + C / 2);
    pos[k + A] = (int32_t)(-sy + C / 2);
    centres[(size_t)(-sy + C / 2) * C + (size_t)(sx + C / 2)] = (float)k;  // coordI().cpt2D()
  }

  mtd_ctx* ctx = nullptr;
  if (mtd_ctx_create(0, &ctx) != MTD_OK) {
    std::cerr << "no GPU available\n";
    return 1;
  }
  Clock t_gpu;
  std::vector<float> fields((size_t)A * C * C * 2);
  {
    void *d_obj = nullptr, *d_pos = nullptr, *d_spec = nullptr, *d_f = nullptr;
    hip_or_die(hipMalloc(&d_obj, nv * 4), "hipMalloc");
    hip_or_die(hipMalloc(&d_pos, pos.size() * 4), "hipMalloc");
    hip_or_die(hipMalloc(&d_spec, fields.size() * 4), "hipMalloc");
    hip_or_die(hipMalloc(&d_f, fields.size() * 4), "hipMalloc");
    hip_or_die(hipMemcpy(d_obj, obj.data(), nv * 4, hipMemcpyHostToDevice), "upload");
    hip_or_die(hipMemcpy(d_pos, pos.data(), pos.size() * 4, hipMemcpyHostToDevice), "upload");
    mtd_or_die(ctx, mtd_forward_spectra(ctx, &g, (const float*)d_obj, nullptr,
                                        (const int32_t*)d_pos, A, (float*)d_spec),
               "forward model");
    mtd_or_die(ctx, mtd_spectra_to_fields(ctx, &g, (const float*)d_spec, (const int32_t*)d_pos, A,
                                          (float*)d_f),
               "field synthesis");
    mtd_or_die(ctx, mtd_ctx_sync(ctx), "sync");
    hip_or_die(hipMemcpy(fields.data(), d_f, fields.size() * 4, hipMemcpyDeviceToHost),
               "download");
    for (void* p : {d_obj, d_pos, d_spec, d_f}) (void)hipFree(p);
  }
  mtd_ctx_destroy(ctx);
  std::cout << "extraction hologramme (GPU)= " << t_gpu.s() * 1e3 << " ms\n";

  const std::string dimImg = std::to_string(C) + "x" + std::to_string(C) + "x" + std::to_string(A);
  const size_t np = (size_t)A * C * C;
  std::vector<double> re(np), im(np);
  for (size_t i = 0; i < np; ++i) {
    re[i] = fields[2 * i];
    im[i] = fields[2 * i + 1];
  }
  bool ok = mtd::write_raw(chemin_result + "/UBornfinal_Re" + dimImg + ".raw", re.data(), np * 8,
                           false);
  ok &= mtd::write_raw(chemin_result + "/UBornfinal_Im" + dimImg + ".raw", im.data(), np * 8,
                       false);
  std::vector<double> tab(2 * (size_t)A);
  for (size_t i = 0; i < tab.size(); ++i) tab[i] = pos[i];
  ok &= mtd::write_raw(chemin_result + "/tab_posSpec.raw", tab.data(), tab.size() * 8, false);
  const double param[5] = {(double)nxmax, (double)A, rayon, (double)dimROI, tp_holo};
  ok &= mtd::write_raw(chemin_result + "/parametres.raw", param, sizeof param, false);
  ok &= mtd::write_tiff_f32(chemin_result + "/obj_Re.tif", obj.data(), V, V, V, tp_tomo, 4);
  // main.cpp:199: the path is joined without a separator
  ok &= mtd::write_tiff_f32(chemin_result + "centres_simul.tif", centres.data(), C, C, 1, tp_holo,
                            1);
  if (!ok) {
    std::cerr << "Erreur d'écriture dans " << chemin_result << "\n";
    return 1;
  }
  std::cout << "Résultats écrits dans " << chemin_result << "\n";
  return 0;
}

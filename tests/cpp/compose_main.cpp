// A main-style program over the C++ mirror of gpuvmem's plugin interfaces:
// the composition of src/main.cu:147-203 (globals, factory-created terms,
// ObjectiveFunction, an optimiser by id) with the device set-up replaced by
// gvm_create / gvm_add_visibilities, as a maintainer would write it.
//
// Input (argv[1], little-endian): gvm_geometry; noise float[M*N]; int nslices;
// per slice: float freq, long nvis, double uvw[3*nvis] (wavelengths, u <= 0),
// float vo[2*nvis], float weight[nvis].  argv[2]: optimiser id, argv[3]:
// iterations; argv[4] (optional) "float-steps": the image mapping is filled the
// reference's way, with the float-step newP / evaluateXt only.  Prints
// "trajectory" and the objective after each iteration.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gvm_hip.h"
#include "gvm/factory.h"
#include "gvm/framework.h"
#include "gvm/image.h"
#include "gvm/objectivefunction.h"
#include "gvm/optimizers.h"
#include "gvm/terms.h"

template <class T>
static void read_into(FILE* f, T* dst, size_t count) {
  if (std::fread(dst, sizeof(T), count, f) != count) {
    std::fprintf(stderr, "short input\n");
    std::exit(2);
  }
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s INPUT OPTIMIZER ITERATIONS [float-steps]\n", argv[0]);
    return 2;
  }
  const bool float_steps = argc > 4 && std::string(argv[4]) == "float-steps";
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  gvm_geometry geo;
  read_into(f, &geo, 1);
  std::vector<float> noise((size_t)geo.M * geo.N);
  read_into(f, noise.data(), noise.size());
  int nsl = 0;
  read_into(f, &nsl, 1);

  try {
    // set-up: replaces MFS::setDevice
    gvm_ctx* ctx = nullptr;
    gvm::check(gvm_create(&geo, noise.data(), &ctx), "gvm_create");
    for (int s = 0; s < nsl; s++) {
      float freq = 0.f;
      long nvis = 0;
      read_into(f, &freq, 1);
      read_into(f, &nvis, 1);
      std::vector<double> uvw(3 * (size_t)nvis);
      std::vector<float> vo(2 * (size_t)nvis), w((size_t)nvis);
      read_into(f, uvw.data(), uvw.size());
      read_into(f, vo.data(), vo.size());
      read_into(f, w.data(), w.size());
      int slot = 0;
      gvm::check(gvm_add_visibilities(ctx, freq, nvis, uvw.data(), vo.data(), w.data(), &slot),
                 "gvm_add_visibilities");
    }
    std::fclose(f);

    // the reference's globals
    auto& g = gvm::globals();
    g.ctx = ctx;
    g.M = geo.M;
    g.N = geo.N;
    g.image_count = 2;
    g.penalizators = {0.01f};  // the -Z list: lambda of the entropy term
    g.eta = -1.0f;
    g.initial_values = {geo.minpix, 0.0f};
    g.flag_opt = 0;

    // composition by factory id, Chi2 first
    ObjectiveFunction* of = createObject<ObjectiveFunction, std::string>("ObjectiveFunction");
    Fi* chi2 = createObject<Fi, std::string>("Chi2");
    Fi* entropy = createObject<Fi, std::string>("Entropy");
    chi2->configure(-1, 0, 0, false);
    entropy->configure(0, 0, 0, false);
    entropy->setPrior(0.001f);
    // Fi::setS / setDS (include/classes/fi.cuh:44-51): the term takes over
    // buffers its caller allocated
    const size_t n = (size_t)geo.M * geo.N;
    void* s_buf = nullptr;
    void* ds_buf = nullptr;
    gvm::check(gvm_malloc(&s_buf, sizeof(float) * n), "gvm_malloc");
    gvm::check(gvm_malloc(&ds_buf, sizeof(float) * n), "gvm_malloc");
    gvm::check(gvm_vec_zero(ctx, (float*)s_buf, (long)n), "gvm_vec_zero");
    gvm::check(gvm_vec_zero(ctx, (float*)ds_buf, (long)n), "gvm_vec_zero");
    entropy->setS((float*)s_buf);
    entropy->setDS((float*)ds_buf);
    // Fi::setInu / getInu (fi.cuh:42-43): a device I_nu buffer kept for the caller
    void* inu = nullptr;
    gvm::check(gvm_malloc(&inu, sizeof(gvm_complex) * n), "gvm_malloc");
    chi2->setInu((gvm_complex*)inu);
    if (chi2->getInu() != (gvm_complex*)inu) {
      std::fprintf(stderr, "getInu does not return the buffer given to setInu\n");
      return 3;
    }
    of->addFi(chi2);
    of->addFi(entropy);
    of->configure(geo.N, geo.M, 2);

    // the image on the device: I_nu_0 = MINPIX, alpha = 0 (src/mfs.cu:737-745)
    std::vector<float> host(2 * n, 0.0f);
    for (size_t k = 0; k < n; k++) host[k] = geo.minpix;
    void* dev = nullptr;
    gvm::check(gvm_malloc(&dev, sizeof(float) * 2 * n), "gvm_malloc");
    gvm::check(gvm_memcpy_h2d(dev, host.data(), sizeof(float) * 2 * n), "gvm_memcpy_h2d");
    Image image((float*)dev, 2);
    imageMap* mapping = nullptr;
    if (float_steps) {
      // src/mfs.cu:794-811: positivity on I_nu_0, none on alpha
      mapping = new imageMap[2];
      mapping[0].newP = particularNewP;
      mapping[0].evaluateXt = particularEvaluateXt;
      mapping[1].newP = defaultNewP;
      mapping[1].evaluateXt = defaultEvaluateXt;
    } else {
      mapping = makeFunctionMapping(2, false);
    }
    image.setFunctionMapping(mapping);

    Optimizer* opt = createObject<Optimizer, std::string>(std::string(argv[2]));
    opt->setObjectiveFunction(of);
    opt->setImage(&image);
    opt->setTotalIterations(std::atoi(argv[3]));
    opt->setK(10);
    opt->setFlag(0);
    opt->optimize();

    std::printf("trajectory");
    for (double v : opt->getTrajectory()) std::printf(" %.17g", v);
    std::printf("\n");

    delete opt;
    delete[] mapping;
    delete chi2;
    delete entropy;  // frees s_buf and ds_buf
    delete of;
    gvm_free(inu);
    gvm_free(dev);
    gvm_destroy(ctx);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

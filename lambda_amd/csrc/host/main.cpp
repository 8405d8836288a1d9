// lambda CLI: ./lambda <scene.nls> [-s SPP -b MAX_BOUNCES -p BOUNCE_PROB -Y MID_GREY]
// Same flags, defaults and outputs as the reference's src/main.cpp:13-47.
#include <unistd.h>

#include <cstdio>
#include <stdexcept>
#include <string>

#include "renderer.hpp"
#include "scene.hpp"

int main(int argc, char **argv) {
  lamh::Renderer renderer;
  float Y = 0.12f;
  int opt;
  try {
    while ((opt = getopt(argc, argv, "b:s:p:Y:")) != -1) {
      switch (opt) {
        case 'b': renderer.MAX_SCATTERINGS = std::stoi(optarg); break;
        case 's': renderer.SPP = std::stoi(optarg); break;
        case 'p': renderer.SAMPLE_P = std::stof(optarg); break;
        case 'Y': Y = std::stof(optarg); break;
        default: std::puts("Usage: ./lambda <fileName> [args]"); return 1;
      }
    }
  } catch (const std::exception &) {
    std::puts("Usage: ./lambda <fileName> [args]");
    return 1;
  }
  if (optind >= argc) { std::puts("Usage: ./lambda <fileName> [args]"); return 1; }
  const char *fileName = argv[optind];
  lamh::Scene scene;
  std::string err;
  if (lamh::load_nls(fileName, scene, err) != LAM_OK) {
    std::fprintf(stderr, "%s\n", err.c_str());
    std::puts("Failed to load scene.");
    return 1;
  }
  try {
    lamh::rendering raw, img;
    renderer.render(scene, raw);
    renderer.saveImage(raw, fileName, "-raw");
    renderer.toneMap(raw, img, Y);
    renderer.saveImage(img, fileName);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    std::puts("Failed to render scene.");
    return 1;
  }
  return 0;
}

// d2 command line on the MI355X engine: the reference's flags and modes
// (src/app/main.cc:53-282, src/app/README.md) over the reference-compatible
// ABI of libd2mi.so.  Modes: clustering (default), --load <c.d2> (start from
// given centroids), --eval <c.d2> (assign only), --prepare_batches <k> [-Q]
// (write shards and their index file, no device work).
#include <getopt.h>

#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "d2mi_clustering.h"

static std::vector<std::string> split(const std::string &s, char delim) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, delim)) out.push_back(item);
  return out;
}

int main(int argc, char *argv[]) {
  int size_of_phases = 1;
  long size_of_samples = 0;
  char *ss1 = nullptr, *ss2 = nullptr, *ss3 = nullptr, *filename = nullptr,
       *centroid_filename = nullptr, *meta_filename = nullptr, *output_filename = nullptr;
  bool is_eval = false, is_load = false, is_pre_processed = false;
  char use_triangle = 1;
  int selected_phase = -1, number_of_clusters = 3, max_iters = 100, num_of_batches = 0;
  static struct option long_options[] = {
      {"strides", 1, 0, 's'},   {"phase", 1, 0, 'p'},      {"ifile", 1, 0, 'i'},
      {"ofile", 1, 0, 'o'},     {"metafile", 1, 0, 'D'},   {"phase_only", 1, 0, 't'},
      {"clusters", 1, 0, 'k'},  {"max_iters", 1, 0, 'm'},  {"centroid_method", 1, 0, 'M'},
      {"non_triangle", 0, 0, 'T'}, {"prepare_batches", 1, 0, 'P'}, {"pre_process", 0, 0, 'Q'},
      {"types", 1, 0, 'E'},     {"eval", 1, 0, 'e'},       {"load", 1, 0, 'L'},
      {nullptr, 0, nullptr, 0}};
  int ch, idx = 0;
  while ((ch = getopt_long(argc, argv, "p:n:s:i:o:D:d:t:k:m:M:TQP:E:e:L:", long_options, &idx)) !=
         -1) {
    switch (ch) {
      case 'i': filename = optarg; break;
      case 'D': meta_filename = optarg; break;
      case 'o': output_filename = optarg; break;
      case 'e': centroid_filename = optarg; is_eval = true; break;
      case 'L': centroid_filename = optarg; is_load = true; break;
      case 'p': size_of_phases = atoi(optarg); break;
      case 'n': size_of_samples = atol(optarg); break;
      case 'd': ss1 = optarg; break;
      case 's': ss2 = optarg; break;
      case 'E': ss3 = optarg; break;
      case 't': selected_phase = atoi(optarg); break;
      case 'k': number_of_clusters = atoi(optarg); break;
      case 'm': max_iters = atoi(optarg); break;
      case 'M': d2_alg_type = atoi(optarg); break;
      case 'T': use_triangle = 0; break;
      case 'Q': is_pre_processed = true; break;
      case 'P': num_of_batches = atoi(optarg); break;
      default:
        fprintf(stderr, "?? getopt returned character code 0%o ??\n", ch);
        return 1;
    }
  }
  if (!filename || !ss2 || size_of_samples <= 0 || size_of_phases <= 0 ||
      number_of_clusters <= 0 || (is_eval && is_load)) {
    fprintf(stderr, "usage: %s -i <file.d2> -n <objects> -d <dims> -s <strides> [-p phases] "
            "[-E types] [-k K] [-m max_iters] [-t phase] [-T] [-o out] [--load c.d2 | --eval c.d2]\n",
            argv[0]);
    return 1;
  }
  if (d2_alg_type != D2_CENTROID_BADMM) {
    fprintf(stderr, "-M %d: only the B-ADMM centroid method (0) runs on the device\n", d2_alg_type);
    return 1;
  }
  std::vector<std::string> s1 = ss1 ? split(ss1, ',') : std::vector<std::string>(size_of_phases, "0");
  std::vector<std::string> s2 = split(ss2, ',');
  std::vector<std::string> s3 = ss3 ? split(ss3, ',') : std::vector<std::string>(size_of_phases, "0");
  if ((int)s1.size() != size_of_phases || (int)s2.size() != size_of_phases ||
      (int)s3.size() != size_of_phases || (size_of_phases > 1 && meta_filename)) {
    fprintf(stderr, "-d, -s and -E need one entry per phase; -D needs one phase\n");
    return 1;
  }
  std::vector<int> dims(size_of_phases), strides(size_of_phases), types(size_of_phases);
  std::cout << "Task: " << std::endl;
  for (int i = 0; i < size_of_phases; ++i) {
    dims[i] = atoi(s1[i].c_str());
    strides[i] = atoi(s2[i].c_str());
    types[i] = atoi(s3[i].c_str());
    std::cout << "\t" << i << "-th phase is of type " << types[i] << std::endl;
  }
  mph data;
  if (d2_allocate(&data, size_of_phases, size_of_samples, strides.data(), dims.data(),
                  types.data()) ||
      d2_read(filename, meta_filename, &data)) {
    std::cerr << "Reading " << filename << " failed" << std::endl;
    return 1;
  }
  if (num_of_batches == 0 && is_pre_processed) num_of_batches = 1;  // main.cc:181
  if (num_of_batches > 0) {  // main.cc:185-194
    const int rc = d2_write_split(filename, &data, num_of_batches, is_pre_processed);
    d2_free(&data);
    std::cout << "[Finish!]" << std::endl;
    return rc ? 1 : 0;
  }
  mph c;
  c.ph = nullptr;
  std::string name;
  if (!is_eval) {
    srand((unsigned)time(nullptr));
    std::stringstream hs;
    hs << rand() % 1000000;  // main.cc:216-222
    name = std::string(filename) + "_" + hs.str();
    if (is_load) {
      data.num_of_labels = number_of_clusters;
      d2_init_centroid(&data, &c, selected_phase, 1);
      if (d2_read(centroid_filename, meta_filename, &c)) return 1;
    }
    if (number_of_clusters == 1) p_badmm_options = &badmm_cen_options;  // main.cc:232-236
    if (d2_clustering(number_of_clusters, max_iters, &data, &c, selected_phase, use_triangle,
                      name.c_str()))
      return 1;
    if (output_filename) name = output_filename;
    d2_write((name + "_c.d2").c_str(), &c);
  } else {
    if (d2_assignment(number_of_clusters, &data, &c, selected_phase, centroid_filename,
                      meta_filename))
      return 1;
    std::string cf(centroid_filename);
    cf = cf.substr(0, cf.rfind("_c"));
    const size_t r = cf.rfind("_");
    name = r == std::string::npos ? std::string(filename) : std::string(filename) + cf.substr(r);
  }
  if (output_filename) name = output_filename;
  d2_write_labels((name + ".label").c_str(), &data);
  d2_write_labels_serial((std::string(filename) + ".ind").c_str(), name.c_str(), &data);
  d2_free(&data);
  d2_free(&c);
  std::cout << "[Finish!]" << std::endl;
  return 0;
}

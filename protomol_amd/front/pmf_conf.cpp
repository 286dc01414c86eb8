// pmf_conf.cpp -- the .conf file: keywords, the Integrator block and the
// force definitions inside it.
//
// io/ConfigurationReader.cpp:42-120 reads the file as one stream of
// whitespace-separated words with '#' comments removed; each keyword takes
// the words its value type reads (three reals for a coordinate, the
// balanced { ... } block for the integrator, one word otherwise).
// frontend/IntegratorFactoryDetails.cpp:35-75 splits the block into
// "level <n> <integrator> { <parameters> force ... }" levels, everything
// from the first "force" word on being the forces; ValueTraits<Force>::read
// (base/ValueType.h:1058-1101) cuts the forces at each "force" word and
// after each word that ends with a comma.
#include "pmf.h"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace pmf {

namespace {

std::string lower(const std::string& s) {
  std::string u(s);
  for (auto& ch : u) ch = (char)std::tolower((unsigned char)ch);
  return u;
}

// The keywords this front end reads or accepts (the reference's
// InputValue definitions, frontend/inputValueDefinitions.cpp, and the
// output files of frontend/Output*.cpp); a word that is no keyword is
// skipped, as ConfigurationReader::read skips it with "Ignoring".
bool is_known_key(const std::string& k) {
  static const char* known[] = {
      "firststep", "numsteps", "exclude", "posfile", "coords", "coordinates", "psffile", "structure", "parfile",
      "parameters", "velfile", "temperature", "seed", "outputfreq", "output", "debug", "screen", "allenergiesfile",
      "allenergiesfileoutputfreq", "doallenergiesfile", "finxyzposfile", "finxyzvelfile", "finpdbposfile",
      "finxyzposfileminimalimage", "finpdbposfileminimalimage", "dcdfile", "dcdfileoutputfreq",
      "dcdfileminimalimage", "dodcdfile", "xyzposfile", "xyzposfileoutputfreq", "xyzvelfile", "xyzvelfileoutputfreq",
      "boundaryconditions", "cellbasisvector1", "cellbasisvector2", "cellbasisvector3", "cellorigin", "cellmanager",
      "cellsize", "virialcalc", "molvirialcalc", "removelinearmomentum", "commotion", "removeangularmomentum",
      "angularmomentum", "shake", "shakeepsilon", "shakemaxiter", "rattle", "rattleepsilon", "rattlemaxiter",
      "dihedralmultpsf", "reducedimage", "minimalimage", "diffusionfile", "diffusionfileoutputfreq", "momentumfile",
      "momentumfileoutputfreq", "tempfile", "tempfileoutputfreq", "shadowenergy", "shadowfreq", "shadoworder",
      "checkpointfreq", "checkpointfile", "eigfile", "eigvalfile", "screenoutputfreq", "precision",
      "remtemperaturefile", "numswitches", "eqswitches", "tempspace", "remexchangeratesfile",
      "remexchangeratesfileoutputfreq", "remhistoryfile", "remhistoryfileoutputfreq"};
  for (const char* x : known)
    if (k == x) return true;
  return false;
}

// keywords whose value is a coordinate (three words)
bool is_vector_key(const std::string& k) {
  return k == "cellbasisvector1" || k == "cellbasisvector2" || k == "cellbasisvector3" || k == "cellorigin";
}

void set_key(std::vector<std::pair<std::string, std::string>>& keys, const std::string& k, const std::string& v) {
  for (auto& kv : keys)
    if (kv.first == k) {
      kv.second = v;
      return;
    }
  keys.emplace_back(k, v);
}

bool parse_levels(const std::vector<std::string>& w, std::vector<LevelSpec>& levels, std::string& err) {
  size_t i = 0;
  while (i < w.size()) {
    if (i + 3 >= w.size() || lower(w[i]) != "level" || w[i + 3] != "{") {
      err = "integration definition mismatch, expecting 'level <number> <integrator> { ... }'";
      return false;
    }
    LevelSpec L;
    if (!to_int(w[i + 1], L.level) || L.level < 0) {
      err = "integration level '" + w[i + 1] + "' is not a number";
      return false;
    }
    L.integrator = w[i + 2];
    i += 4;
    std::vector<std::string> params, forces;
    while (i < w.size() && w[i] != "}") {
      if (lower(w[i]) == "force" || !forces.empty())
        forces.push_back(w[i]);
      else
        params.push_back(w[i]);
      i++;
    }
    if (i >= w.size()) {
      err = "integrator level " + std::to_string(L.level) + " is not closed by '}'";
      return false;
    }
    i++;
    for (size_t p = 0; p + 1 < params.size(); p += 2) L.params.emplace_back(lower(params[p]), params[p + 1]);
    if (params.size() % 2) {
      err = "integrator level " + std::to_string(L.level) + ": parameter '" + params.back() + "' has no value";
      return false;
    }
    // forces: cut at "force" and after a word ending in ','
    ForceSpec cur;
    auto flush = [&]() {
      if (!cur.id.empty()) L.forces.push_back(cur);
      cur = ForceSpec();
    };
    for (size_t k = 0; k < forces.size(); k++) {
      std::string f = forces[k];
      if (lower(f) == "force") {
        flush();
        continue;
      }
      bool cut = false;
      if (!f.empty() && f.back() == ',') {
        f.pop_back();
        cut = true;
      }
      if (!f.empty()) cur.id += (cur.id.empty() ? "" : " ") + f;
      if (cut) flush();
    }
    flush();
    // names are the words before the first option, options "-word value..."
    for (ForceSpec& fs : L.forces) {
      std::istringstream ss(fs.id);
      std::string t;
      std::vector<std::string> words;
      while (ss >> t) words.push_back(t);
      size_t k = 0;
      while (k < words.size() && words[k][0] != '-') fs.names.push_back(words[k++]);
      auto is_option = [](const std::string& x) {
        return x.size() > 1 && x[0] == '-' && !std::isdigit((unsigned char)x[1]) && x[1] != '.';
      };
      while (k < words.size()) {
        std::string key = words[k++], val;
        while (k < words.size() && !is_option(words[k])) val += (val.empty() ? "" : " ") + words[k++];
        fs.options.emplace_back(key, val);
      }
    }
    levels.push_back(L);
  }
  return true;
}

}  // namespace

bool Config::has(const std::string& k) const {
  const std::string lk = lower(k);
  for (const auto& kv : keys)
    if (kv.first == lk) return true;
  return false;
}

std::string Config::get(const std::string& k, const std::string& def) const {
  const std::string lk = lower(k);
  for (const auto& kv : keys)
    if (kv.first == lk) return kv.second;
  return def;
}

double Config::real(const std::string& k, double def) const {
  double v;
  return has(k) && to_real(get(k), v) ? v : def;
}

long Config::integer(const std::string& k, long def) const {
  int v;
  return has(k) && to_int(get(k), v) ? v : def;
}

bool Config::boolean(const std::string& k, bool def) const {
  if (!has(k)) return def;
  const std::string v = lower(get(k));
  if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
  if (v == "false" || v == "no" || v == "off" || v == "0") return false;
  return def;
}

// file names are relative to the working directory, as the reference
// opens them
std::string Config::path(const std::string& k) const { return get(k); }

bool read_conf(const std::string& path, Config& cfg, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open configuration file '" + path + "'";
    return false;
  }
  cfg = Config();
  const size_t slash = path.find_last_of('/');
  cfg.dir = slash == std::string::npos ? "" : path.substr(0, slash);
  std::vector<std::string> w;
  std::string line, t;
  while (std::getline(f, line)) {
    std::istringstream ss(line.substr(0, line.find('#')));
    while (ss >> t) w.push_back(t);
  }
  for (size_t i = 0; i < w.size();) {
    const std::string k = lower(w[i++]);
    if (k == "integrator") {
      if (i >= w.size() || w[i] != "{") {
        err = "Integrator must be followed by a { ... } block";
        return false;
      }
      int depth = 0;
      std::vector<std::string> block;
      for (; i < w.size(); i++) {
        if (w[i] == "{") depth++;
        if (w[i] == "}") depth--;
        if (depth == 0) break;
        if (!(depth == 1 && w[i] == "{" && block.empty())) block.push_back(w[i]);
      }
      if (i >= w.size()) {
        err = "the Integrator block is not closed by '}'";
        return false;
      }
      i++;
      if (!parse_levels(block, cfg.levels, err)) return false;
      continue;
    }
    if (!is_known_key(k)) {
      cfg.ignored.push_back(w[i - 1]);
      continue;
    }
    const size_t nw = is_vector_key(k) ? 3 : 1;
    if (i + nw > w.size()) {
      err = "keyword '" + k + "' has no value";
      return false;
    }
    std::string v;
    for (size_t j = 0; j < nw; j++) v += (j ? " " : "") + w[i + j];
    i += nw;
    set_key(cfg.keys, k, v);
  }
  return true;
}

bool apply_overrides(Config& cfg, int argc, const char* const* argv, std::string& err) {
  for (int i = 0; i < argc; i++) {
    std::string a = argv[i];
    if (a.compare(0, 2, "--") != 0) {
      err = "unexpected argument '" + a + "' (use --keyword value)";
      return false;
    }
    const std::string k = lower(a.substr(2));
    const int nw = is_vector_key(k) ? 3 : 1;
    if (i + nw >= argc) {
      err = "option --" + k + " has no value";
      return false;
    }
    std::string v;
    for (int j = 1; j <= nw; j++) v += std::string(j > 1 ? " " : "") + argv[i + j];
    i += nw;
    set_key(cfg.keys, k, v);
  }
  return true;
}

}  // namespace pmf

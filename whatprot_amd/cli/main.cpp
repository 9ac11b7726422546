// whatprot-amd: the command line of whatprot's scoring path over the C ABI of
// libwhatprot_amd (main/main.cc:56-447 and the run_* drivers).
//
//   whatprot-amd classify hmm    -P seqparams.json -S dyeseqs.tsv -R rads.tsv -Y out.csv [-p cutoff]
//   whatprot-amd classify nn     -P seqparams.json -k K -s SIG -T dyetracks.tsv -R rads.tsv -Y out.csv
//   whatprot-amd classify hybrid -P ... -k K -s SIG -H H -S ... -T ... -R ... -Y ... [-p cutoff]
//   whatprot-amd fit -P seqparams.json -x DYESEQ -L THRESHOLD -R rads.tsv [-F fit.json] [-M minutes]
//                    [-b ROUNDS -c CI [-Y bootstrap.csv]]
//   whatprot-amd simulate rad -P ... -t T -g N -S dyeseqs.tsv -R rads.tsv -Y true_ids.tsv
//   whatprot-amd simulate dt  -P ... -t T -g N -S dyeseqs.tsv -T dyetracks.tsv
//
// Flags keep the reference's short and long names and required sets. Files
// are read and written in the reference's formats (io/*.cc). Additions:
// --precision fp64|fp32 for the HMM paths (default fp64, the reference's
// arithmetic) and --seed for simulation and bootstrap resampling (the
// reference seeds from the clock).
#include <chrono>
#include <cctype>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "whatprot_amd.h"

namespace {

// ---- minimal JSON (objects, arrays, numbers, strings, literals) ------------

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
    const Json* get(const std::string& k) const {
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JsonParser {
    const std::string& s;
    size_t i = 0;
    explicit JsonParser(const std::string& text) : s(text) {}
    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("JSON parse error: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    Json value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        Json j;
        if (c == '{') {
            j.kind = Json::Object;
            i++;
            ws();
            if (s[i] == '}') {
                i++;
                return j;
            }
            while (true) {
                ws();
                Json key = value();
                if (key.kind != Json::String) fail("object key");
                ws();
                if (s[i] != ':') fail("':'");
                i++;
                j.obj.emplace_back(key.str, value());
                ws();
                if (s[i] == ',') {
                    i++;
                    continue;
                }
                if (s[i] == '}') {
                    i++;
                    return j;
                }
                fail("',' or '}'");
            }
        }
        if (c == '[') {
            j.kind = Json::Array;
            i++;
            ws();
            if (s[i] == ']') {
                i++;
                return j;
            }
            while (true) {
                j.arr.push_back(value());
                ws();
                if (s[i] == ',') {
                    i++;
                    continue;
                }
                if (s[i] == ']') {
                    i++;
                    return j;
                }
                fail("',' or ']'");
            }
        }
        if (c == '"') {
            j.kind = Json::String;
            i++;
            while (i < s.size() && s[i] != '"') {
                if (s[i] == '\\' && i + 1 < s.size()) i++;
                j.str += s[i++];
            }
            i++;
            return j;
        }
        if (s.compare(i, 4, "true") == 0) {
            j.kind = Json::Bool;
            j.b = true;
            i += 4;
            return j;
        }
        if (s.compare(i, 5, "false") == 0) {
            j.kind = Json::Bool;
            i += 5;
            return j;
        }
        if (s.compare(i, 4, "null") == 0) {
            i += 4;
            return j;
        }
        char* end = nullptr;
        j.kind = Json::Number;
        j.num = std::strtod(s.c_str() + i, &end);
        if (end == s.c_str() + i) fail("value");
        i = (size_t)(end - s.c_str());
        return j;
    }
};

Json read_json(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    JsonParser p(text);
    return p.value();
}

double num(const Json& j, const char* key) {
    const Json* v = j.get(key);
    if (!v || v->kind != Json::Number) throw std::runtime_error(std::string("missing number: ") + key);
    return v->num;
}

// SequencingModel(filename) (sequencing-model.cc:51-89).
wp_model read_seq_model(const std::string& path) {
    Json j = read_json(path);
    wp_model m;
    std::memset(&m, 0, sizeof(m));
    m.p_edman_failure = num(j, "p_edman_failure");
    m.p_detach = num(j, "p_detach");
    m.p_initial_detach = num(j, "p_initial_detach");
    m.p_initial_detach_decay = num(j, "p_initial_detach_decay");
    m.p_initial_block = num(j, "p_initial_block");
    m.p_cyclic_block = num(j, "p_cyclic_block");
    const Json* chans = j.get("channel_models");
    if (!chans || chans->kind != Json::Array) throw std::runtime_error("missing channel_models");
    int C = (int)chans->arr.size();
    if (C > WP_MAX_CHANNELS) throw std::runtime_error("too many channels");
    m.num_channels = C;
    for (int c = 0; c < C; c++) {
        const Json& cd = chans->arr[c];
        wp_channel_model& ch = m.channels[c];
        ch.p_bleach = num(cd, "p_bleach");
        ch.p_dud = num(cd, "p_dud");
        ch.bg_sig = num(cd, "bg_sig");
        ch.mu = num(cd, "mu");
        ch.sig = num(cd, "sig");
        for (int o = 0; o < WP_MAX_CHANNELS; o++) ch.interactions[o] = ch.flat_interactions[o] = 1.0;
        const Json* inter = cd.get("interactions");
        if (inter && inter->kind == Json::Array) {
            for (const Json& it : inter->arr) {
                int other = (int)num(it, "other_channel");
                if (other < 0 || other >= C) throw std::runtime_error("bad other_channel");
                if (const Json* e = it.get("effect")) ch.interactions[other] = 1.0 - e->num;
                if (const Json* e = it.get("flat_effect")) ch.flat_interactions[other] = 1.0 - e->num;
            }
        }
    }
    return m;
}

// FitSettings(num_channels, filename) (fit-settings.cc:44-100).
wp_fit_settings read_fit_settings(const std::string& path, int C) {
    wp_fit_settings fs;
    std::memset(&fs, 0, sizeof(fs));
    if (path.empty()) return fs;
    Json j = read_json(path);
    auto flag = [&](const Json& o, const char* k) {
        const Json* v = o.get(k);
        return v && v->kind == Json::Bool && v->b ? 1 : 0;
    };
    fs.hold_p_edman_failure = flag(j, "hold_p_edman_failure");
    fs.hold_p_detach = flag(j, "hold_p_detach");
    fs.hold_p_initial_detach = flag(j, "hold_p_initial_detach");
    fs.hold_p_initial_detach_decay = flag(j, "hold_p_initial_detach_decay");
    fs.hold_p_initial_block = flag(j, "hold_p_initial_block");
    fs.hold_p_cyclic_block = flag(j, "hold_p_cyclic_block");
    if (const Json* cs = j.get("channel_settings")) {
        for (size_t c = 0; c < cs->arr.size() && (int)c < C; c++) {
            fs.hold_p_bleach[c] = flag(cs->arr[c], "hold_p_bleach");
            fs.hold_p_dud[c] = flag(cs->arr[c], "hold_p_dud");
        }
    }
    return fs;
}

struct DyeSeqs {
    int C = 0;
    std::vector<std::string> strings;
    std::vector<int> counts, ids;
};

// read_dye_seqs (dye-seqs-io.cc:70-87): C, n, then (string, count, id) rows.
DyeSeqs read_dye_seqs(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    DyeSeqs d;
    int n;
    f >> d.C >> n;
    for (int i = 0; i < n; i++) {
        std::string s;
        int c, id;
        f >> s >> c >> id;
        d.strings.push_back(s);
        d.counts.push_back(c);
        d.ids.push_back(id);
    }
    return d;
}

void check(int rc) {
    if (rc != WP_OK) throw std::runtime_error(std::string("whatprot_amd: ") + wp_last_error());
}

struct Rads {
    int T = 0, C = 0, N = 0;
    std::vector<double> v;
};

// read_radiometries (radiometries-io.cc:33-56) through the library's parallel
// reader: T, C, N, then N rows of T*C intensities, each divided by its
// channel's mu of the model as read.
Rads read_radiometries(const std::string& path, const wp_model& true_model) {
    wp_radiometries r;
    check(wp_read_radiometries(path.c_str(), &true_model, &r));
    Rads o;
    o.T = r.num_timesteps;
    o.C = r.num_channels;
    o.N = r.num_reads;
    o.v.assign(r.values, r.values + (size_t)o.N * o.T * o.C);
    wp_free_radiometries(&r);
    return o;
}

// read_dye_tracks (dye-tracks-io.cc:32-62) through the library's parallel
// reader; the arrays stay in the library's allocation.
struct Tracks {
    wp_dye_tracks d;
    Tracks() { std::memset(&d, 0, sizeof(d)); }
    Tracks(const Tracks&) = delete;
    Tracks& operator=(const Tracks&) = delete;
    ~Tracks() { wp_free_dye_tracks(&d); }
};

void read_dye_tracks(const std::string& path, Tracks* t) { check(wp_read_dye_tracks(path.c_str(), &t->d)); }

// write_scored_classifications_raw (scored-classifications-io.cc:54-68).
void write_results(const std::string& path, const std::vector<int>& ids, const std::vector<double>& scores) {
    check(wp_write_scored_classifications(path.c_str(), (int)ids.size(), ids.data(), scores.data()));
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::vector<const char*> cstrs(const std::vector<std::string>& v) {
    std::vector<const char*> out;
    for (const auto& s : v) out.push_back(s.c_str());
    return out;
}

struct Args {
    std::vector<std::string> pos;
    std::map<char, std::string> opt;
    std::string precision = "fp64";
    unsigned long long seed = 1;
    bool has(char c) const { return opt.count(c) != 0; }
    const std::string& get(char c) const { return opt.at(c); }
};

const std::map<std::string, char> kLong = {
    {"numbootstrap", 'b'}, {"confidenceinterval", 'c'}, {"numgenerate", 'g'}, {"neighbors", 'k'},
    {"hmmprune", 'p'},     {"sigma", 's'},              {"timesteps", 't'},   {"dyeseqstring", 'x'},
    {"fitsettings", 'F'},  {"passthrough", 'H'},        {"stoppingthreshold", 'L'}, {"maxruntime", 'M'},
    {"seqparams", 'P'},    {"radiometries", 'R'},       {"dyeseqs", 'S'},     {"dyetracks", 'T'},
    {"results", 'Y'}};

const char* kUsage =
    "usage: whatprot-amd MODE [VARIANT] [OPTS...]\n"
    "  classify hmm    -P seqparams -S dyeseqs -R radiometries -Y results [-p hmmprune]\n"
    "  classify nn     -P seqparams -k neighbors -s sigma -T dyetracks -R radiometries -Y results\n"
    "  classify hybrid -P seqparams -k neighbors -s sigma -H passthrough -S dyeseqs -T dyetracks\n"
    "                  -R radiometries -Y results [-p hmmprune]\n"
    "  fit             -P seqparams -x dyeseqstring -L stoppingthreshold -R radiometries\n"
    "                  [-F fitsettings] [-M maxruntime (minutes)]\n"
    "                  [-b numbootstrap -c confidenceinterval [-Y results]]\n"
    "  simulate rad    -P seqparams -t timesteps -g numgenerate -S dyeseqs -R radiometries -Y results\n"
    "  simulate dt     -P seqparams -t timesteps -g numgenerate -S dyeseqs -T dyetracks\n"
    "  extras: --precision fp64|fp32 (HMM arithmetic), --seed N (simulation)\n";

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; i++) {
        std::string s = argv[i];
        if (s == "-h" || s == "--help") {
            std::cout << kUsage;
            std::exit(0);
        }
        if (s == "--precision" && i + 1 < argc) {
            a.precision = argv[++i];
            continue;
        }
        if (s == "--seed" && i + 1 < argc) {
            a.seed = std::strtoull(argv[++i], nullptr, 10);
            continue;
        }
        char key = 0;
        std::string val;
        if (s.size() >= 2 && s[0] == '-' && s[1] == '-') {
            std::string name = s.substr(2);
            size_t eq = name.find('=');
            if (eq != std::string::npos) {
                val = name.substr(eq + 1);
                name = name.substr(0, eq);
            }
            auto it = kLong.find(name);
            if (it == kLong.end()) throw std::runtime_error("unknown option " + s);
            key = it->second;
            if (eq == std::string::npos) {
                if (i + 1 >= argc) throw std::runtime_error("missing value for " + s);
                val = argv[++i];
            }
        } else if (s.size() >= 2 && s[0] == '-' && !std::isdigit((unsigned char)s[1])) {
            key = s[1];
            if (s.size() > 2) {
                val = s.substr(2);
            } else {
                if (i + 1 >= argc) throw std::runtime_error("missing value for " + s);
                val = argv[++i];
            }
        } else {
            a.pos.push_back(s);
            continue;
        }
        a.opt[key] = val;
    }
    return a;
}

// The reference's required-option rule (main.cc:368-447): exactly these
// options, plus any listed as optional.
bool needs(const Args& a, const std::string& req, const std::string& optional) {
    for (char c : req)
        if (!a.has(c)) return false;
    for (const auto& kv : a.opt)
        if (req.find(kv.first) == std::string::npos && optional.find(kv.first) == std::string::npos) return false;
    return true;
}

int precision_of(const Args& a) {
    if (a.precision == "fp64") return WP_PRECISION_FP64;
    if (a.precision == "fp32") return WP_PRECISION_FP32;
    throw std::runtime_error("--precision must be fp64 or fp32");
}

void need_rads_shape(const Rads& r, int C) {
    if (r.C != C) throw std::runtime_error("radiometries and model disagree on the number of channels");
}

int classify_hmm(const Args& a) {
    double t0 = now();
    wp_model truth = read_seq_model(a.get('P')), model;
    wp_model_with_mu_as_one(&truth, &model);
    double cutoff = a.has('p') ? std::stod(a.get('p')) : DBL_MAX;
    DyeSeqs ds = read_dye_seqs(a.get('S'));
    std::cout << "Read " << ds.strings.size() << " dye seqs.\n";
    Rads r = read_radiometries(a.get('R'), truth);
    need_rads_shape(r, model.num_channels);
    std::cout << "Read " << r.N << " radiometries.\n";
    auto strs = cstrs(ds.strings);
    wp_hmm_classifier* clf = nullptr;
    check(wp_hmm_classifier_create(r.T, r.C, &model, cutoff, (int)ds.strings.size(), strs.data(), ds.counts.data(),
                                   ds.ids.data(), precision_of(a), &clf));
    std::vector<int> id(r.N);
    std::vector<double> adj(r.N);
    double t1 = now();
    check(wp_hmm_classify(clf, r.N, r.v.data(), id.data(), adj.data(), nullptr, nullptr));
    std::cout << "Finished classification (" << now() - t1 << " seconds).\n";
    wp_hmm_classifier_destroy(clf);
    write_results(a.get('Y'), id, adj);
    std::cout << "Total run time: " << now() - t0 << " seconds.\n";
    return 0;
}

void load_tracks(const Args& a, int C, Tracks* t) {
    read_dye_tracks(a.get('T'), t);
    if (t->d.num_channels != C) throw std::runtime_error("dye-tracks and model disagree on the number of channels");
    std::cout << "Read " << t->d.num_dye_tracks << " dye tracks.\n";
}

int classify_nn(const Args& a) {
    double t0 = now();
    wp_model truth = read_seq_model(a.get('P')), model;
    wp_model_with_mu_as_one(&truth, &model);
    Tracks tr;
    load_tracks(a, model.num_channels, &tr);
    Rads r = read_radiometries(a.get('R'), truth);
    need_rads_shape(r, model.num_channels);
    wp_nn_classifier* nn = nullptr;
    check(wp_nn_classifier_create(tr.d.num_timesteps, tr.d.num_channels, &model, std::stoi(a.get('k')),
                                  std::stod(a.get('s')), tr.d.num_dye_tracks, tr.d.counts, tr.d.source_offsets,
                                  tr.d.source_ids, tr.d.source_counts, tr.d.source_hits, &nn));
    std::vector<int> id(r.N);
    std::vector<double> sc(r.N), tot(r.N), adj(r.N);
    double t1 = now();
    check(wp_nn_classify(nn, r.N, r.v.data(), id.data(), sc.data(), tot.data()));
    std::cout << "Finished classification (" << now() - t1 << " seconds).\n";
    for (int i = 0; i < r.N; i++) adj[i] = sc[i] / tot[i];
    wp_nn_classifier_destroy(nn);
    write_results(a.get('Y'), id, adj);
    std::cout << "Total run time: " << now() - t0 << " seconds.\n";
    return 0;
}

int classify_hybrid(const Args& a) {
    double t0 = now();
    wp_model truth = read_seq_model(a.get('P')), model;
    wp_model_with_mu_as_one(&truth, &model);
    double cutoff = a.has('p') ? std::stod(a.get('p')) : DBL_MAX;
    DyeSeqs ds = read_dye_seqs(a.get('S'));
    Tracks tr;
    load_tracks(a, model.num_channels, &tr);
    Rads r = read_radiometries(a.get('R'), truth);
    need_rads_shape(r, model.num_channels);
    auto strs = cstrs(ds.strings);
    wp_hybrid_classifier* hc = nullptr;
    check(wp_hybrid_classifier_create(tr.d.num_timesteps, tr.d.num_channels, &model, cutoff, std::stoi(a.get('k')),
                                      std::stod(a.get('s')), tr.d.num_dye_tracks, tr.d.counts, tr.d.source_offsets,
                                      tr.d.source_ids, tr.d.source_counts, tr.d.source_hits, std::stoi(a.get('H')),
                                      (int)ds.strings.size(), strs.data(), ds.counts.data(), ds.ids.data(),
                                      precision_of(a), &hc));
    std::vector<int> id(r.N);
    std::vector<double> adj(r.N);
    double t1 = now();
    check(wp_hybrid_classify(hc, r.N, r.v.data(), id.data(), adj.data()));
    std::cout << "Finished classification (" << now() - t1 << " seconds).\n";
    wp_hybrid_classifier_destroy(hc);
    write_results(a.get('Y'), id, adj);
    std::cout << "Total run time: " << now() - t0 << " seconds.\n";
    return 0;
}

// SequencingModel::debug_string (sequencing-model.cc:168-184).
std::string debug_string(const wp_model& m) {
    auto f = [](double x) { return std::to_string(x); };
    std::string s = "Edman failure rate: " + f(m.p_edman_failure) + ", ";
    s += "Detach rate: " + f(m.p_detach) + ", ";
    s += "Initial detach rate: " + f(m.p_initial_detach) + ", ";
    s += "Initial detach decay rate: " + f(m.p_initial_detach_decay) + ", ";
    s += "Initial blocked N-terminus rate: " + f(m.p_initial_block) + ", ";
    s += "Cyclic blocked N-terminus rate: " + f(m.p_cyclic_block);
    for (int c = 0; c < m.num_channels; c++) {
        s += ", Channel " + std::to_string(c) + " info: (";
        s += "Bleach rate: " + f(m.channels[c].p_bleach) + ", Dud rate: " + f(m.channels[c].p_dud) + ")";
    }
    return s;
}

int fit(const Args& a) {
    double t0 = now();
    wp_model truth = read_seq_model(a.get('P')), model;
    wp_model_with_mu_as_one(&truth, &model);
    int C = model.num_channels;
    wp_fit_settings fs = read_fit_settings(a.has('F') ? a.get('F') : "", C);
    Rads r = read_radiometries(a.get('R'), truth);
    need_rads_shape(r, C);
    std::cout << "Read " << r.N << " radiometries.\n";
    double minutes = a.has('M') ? std::stod(a.get('M')) : (double)INT_MAX;
    wp_hmm_fitter* f = nullptr;
    check(wp_hmm_fitter_create(r.T, r.C, a.get('x').c_str(), r.N, r.v.data(), precision_of(a), &f));
    wp_model out;
    double step = 0.0, log_l = 0.0;
    int iters = 0;
    double t1 = now();
    const double ci = a.has('c') ? std::stod(a.get('c')) : 0.0;
    if (ci == 0.0) {  // run-fit.cc:88-99
        check(wp_hmm_fit(f, &model, &fs, std::stod(a.get('L')), minutes * 60.0, 0, &out, &step, &log_l, &iters));
        std::cout << "Finished parameter fitting (" << now() - t1 << " seconds, " << iters << " iterations).\n";
    } else {  // run-fit.cc:100-130 and bootstrap-fit.cc:37-173
        const int b = std::stoi(a.get('b'));
        std::vector<wp_model> models(std::max(b, 1));
        std::vector<double> log_ls(std::max(b, 1));
        wp_model lo, hi;
        double worst = 0.0;
        check(wp_hmm_fit_bootstrap(f, &model, &fs, std::stod(a.get('L')), minutes * 60.0, b, ci, a.seed,
                                   models.data(), log_ls.data(), &out, &step, &log_l, &lo, &hi, &worst));
        std::cout << "lower bounds: " << debug_string(lo) << "\n";
        std::cout << "upper bounds: " << debug_string(hi) << "\n";
        std::cout << "worst bootstrap step size: " << worst << "\n";
        std::cout << "Finished parameter fitting (" << now() - t1 << " seconds).\n";
        if (a.has('Y')) {  // write_params (params-io.cc:37-54)
            std::ofstream y(a.get('Y'));
            y << "p_edman_failure,p_detach,p_initial_detach,p_initial_detach_decay,p_initial_block,p_cyclic_block";
            for (int c = 0; c < C; c++) y << ",ch" << c << ":p_bleach,ch" << c << ":p_dud";
            y << ",log(L)\n" << std::setprecision(17);
            for (int i = 0; i < b; i++) {
                const wp_model& m = models[i];
                y << m.p_edman_failure << "," << m.p_detach << "," << m.p_initial_detach << ","
                  << m.p_initial_detach_decay << "," << m.p_initial_block << "," << m.p_cyclic_block;
                for (int c = 0; c < C; c++) y << "," << m.channels[c].p_bleach << "," << m.channels[c].p_dud;
                y << "," << log_ls[i] << "\n";
            }
        }
    }
    wp_hmm_fitter_destroy(f);
    // print_parameter_results and print_final_step_size (cmd-line-out.cc)
    std::cout << "Parameters:\n" << debug_string(out) << "\n\n";
    std::streamsize prec = std::cout.precision();
    std::cout << "log(L): " << std::setprecision(17) << log_l << "\n\n" << std::setprecision(prec);
    std::cout << "Final step size: " << step << "\n";
    std::cout << "Total run time: " << now() - t0 << " seconds.\n";
    return 0;
}

int simulate(const Args& a, bool rad) {
    wp_model truth = read_seq_model(a.get('P'));
    DyeSeqs ds = read_dye_seqs(a.get('S'));
    int T = std::stoi(a.get('t')), g = std::stoi(a.get('g'));
    auto strs = cstrs(ds.strings);
    int C = truth.num_channels;
    if (rad) {  // generate_radiometries with the true model (run-simulate-rad.cc)
        std::vector<double> v((size_t)g * T * C);
        std::vector<int> idx(g);
        check(wp_simulate_radiometries_gpu(&truth, T, C, (int)ds.strings.size(), strs.data(), ds.counts.data(), g, a.seed,
                                       v.data(), idx.data()));
        check(wp_write_radiometries(a.get('R').c_str(), nullptr, T, C, g, v.data()));
        std::ofstream y(a.get('Y'));  // write_ys (radiometries-io.cc:104-114)
        y << g << "\n";
        for (int i = 0; i < g; i++) y << ds.ids[idx[i]] << "\n";
        return 0;
    }
    long long total = 0;
    for (int c : ds.counts) total += c;
    int cap = (int)std::min<long long>(INT_MAX / 2, std::max<long long>(1, total * g));
    int n = 0, ns = 0;
    std::vector<short> counts((size_t)cap * T * C);
    std::vector<int> off(cap + 1), ids(cap), cnts(cap), hits(cap);
    check(wp_simulate_dye_tracks_gpu(&truth, T, C, (int)ds.strings.size(), strs.data(), ds.counts.data(), ds.ids.data(), g,
                                     a.seed, cap, cap, &n, &ns, counts.data(), off.data(), ids.data(), cnts.data(),
                                     hits.data()));
    // write_dye_tracks (dye-tracks-io.cc:66-103)
    check(wp_write_dye_tracks(a.get('T').c_str(), T, C, n, counts.data(), off.data(), ids.data(), cnts.data(),
                              hits.data()));
    return 0;
}

int usage() {
    std::cout << "\nINCORRECT USAGE\n\n" << kUsage;
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Args a = parse(argc, argv);
        if (a.pos.empty()) return usage();
        const std::string& mode = a.pos[0];
        if (mode == "classify" && a.pos.size() == 2) {
            const std::string& v = a.pos[1];
            if (v == "hmm" && needs(a, "PSRY", "p")) return classify_hmm(a);
            if (v == "nn" && needs(a, "PksTRY", "")) return classify_nn(a);
            if (v == "hybrid" && needs(a, "PksHSTRY", "p")) return classify_hybrid(a);
            return usage();
        }
        if (mode == "fit" && a.pos.size() == 1) {  // main.cc:413-446: -b and -c both or neither
            if (a.has('b') != a.has('c')) return usage();
            if (a.has('b') ? needs(a, "PLxRbc", "FMY") : needs(a, "PLxR", "FM")) return fit(a);
            return usage();
        }
        if (mode == "simulate" && a.pos.size() == 2) {
            if (a.pos[1] == "rad" && needs(a, "PtgSRY", "")) return simulate(a, true);
            if (a.pos[1] == "dt" && needs(a, "PtgST", "")) return simulate(a, false);
            return usage();
        }
        return usage();
    } catch (const std::exception& e) {
        std::cerr << "whatprot-amd: " << e.what() << "\n";
        return 1;
    }
}

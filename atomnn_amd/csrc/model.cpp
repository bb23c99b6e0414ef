// Parser for AtomNN "ann" potential files and the device-plan builder.
//
// The grammar follows the writer/reader pairs of the reference:
//   NNP::read            nnp/nnp.cpp:244-308
//   Type::read           nnp/type.cpp:32-70
//   BasisR::read         nnp/basis_radial.cpp:123-144
//   BasisA::read/init    nnp/basis_angular.cpp:113-138, 168-173
//   ANN::read            ml/nn.cpp:1349-1425
//   NNH::init_input      nnp/nnh.cpp:95-121   (descriptor offsets)
// Error texts repeat the reference's exception messages where it throws.
#include "model.hpp"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

namespace atomnn {

static const double kPi = 3.141592653589793;  // math/const.hpp:9

// str/string.cpp:17-28
int name_hash(const char* s) {
    const long long p = 31, m = 2147483647;
    long long val = 0, pw = 1;
    for (; *s; ++s) {
        val = (val + (long long)(*s - 'a' + 1) * pw) % m;
        pw = (pw * p) % m;
    }
    return (int)val;
}

int HostModel::index(const std::string& name) const {
    const int h = name_hash(name.c_str());
    for (int i = 0; i < T; ++i)
        if (sp[i].hash == h) return i;
    return -1;
}

namespace {

// Whitespace tokeniser over one line at a time (delimiters " \r\t\n", str/string.hpp:16).
class Lines {
  public:
    explicit Lines(FILE* f) : f_(f) {}
    const std::vector<std::string>& next(const char* what) {
        std::string line;
        int ch;
        while ((ch = std::fgetc(f_)) != EOF) {
            line.push_back((char)ch);
            if (ch == '\n') break;
        }
        if (line.empty()) throw std::invalid_argument(std::string("unexpected end of potential file reading ") + what);
        tok_.clear();
        const char* ws = " \r\t\n";
        size_t p = line.find_first_not_of(ws);
        while (p != std::string::npos) {
            const size_t e = line.find_first_of(ws, p);
            tok_.push_back(line.substr(p, e == std::string::npos ? std::string::npos : e - p));
            p = (e == std::string::npos) ? e : line.find_first_not_of(ws, e);
        }
        what_ = what;
        return tok_;
    }
    const std::string& tok(size_t i) const {
        if (i >= tok_.size()) throw std::invalid_argument(std::string("missing field in potential file line: ") + what_);
        return tok_[i];
    }
    size_t size() const { return tok_.size(); }

  private:
    FILE* f_;
    std::vector<std::string> tok_;
    const char* what_ = "";
};

// basis.cpp:33-46, 160-168
CutoffSpec make_cutoff(const std::string& norm, const std::string& cut, double rc) {
    CutoffSpec c;
    if (rc <= 0) throw std::invalid_argument("Basis(double,cutoff::Name,cutoff::Norm,int): Invalid cutoff radius.");
    if (cut == "STEP") c.kind = CUT_STEP;
    else if (cut == "COS") c.kind = CUT_COS;
    else if (cut == "TANH") c.kind = CUT_TANH;
    else throw std::invalid_argument("Basis:cutf(double): invalid cutoff type.");
    c.rc = rc;
    c.rci = 1.0 / rc;
    c.pirci = kPi / rc;
    if (norm == "UNIT") c.norm = 1.0;
    else if (norm == "VOL") c.norm = 1.0 / (0.5 * 4.0 / 3.0 * (kPi * kPi - 6.0) / kPi * rc * rc * rc);
    else throw std::invalid_argument("Basis::norm(NormT::type,double): invalid normalization scheme.");
    return c;
}

// math/special.cpp:111-120
double powint(double x, int n) {
    if (n == 0) return 1.0;
    if (x == 0.0) return 0.0;
    int nn = n > 0 ? n : -n;
    double ww = x, yy = 1.0;
    for (; nn != 0; nn >>= 1, ww *= ww)
        if (nn & 1) yy *= ww;
    return n > 0 ? yy : 1.0 / yy;
}

void read_radial(Lines& in, RadialSpec& r) {
    in.next("BasisR");
    const std::string phi = in.tok(4);
    const int n = std::atoi(in.tok(5).c_str());
    r.cut = make_cutoff(in.tok(1), in.tok(2), std::atof(in.tok(3).c_str()));
    static const char* names[] = {"GAUSSIAN", "TANH", "SOFTPLUS", "LOGCOSH", "SWISH", "MISH"};
    r.phi = -1;
    for (int k = 0; k < 6; ++k)
        if (phi == names[k]) r.phi = k;
    if (r.phi < 0)  // basis_radial.cpp:70
        throw std::invalid_argument("BasisR(rc,cutoff::Name,cutoff::Norm,int,PhiRN): invalid radial function type");
    if (n < 0) throw std::invalid_argument("Basis::resize(int): invalid number of functions.");
    r.rs.resize(n);
    r.eta.resize(n);
    for (int i = 0; i < n; ++i) {
        in.next("BasisR parameters");
        r.rs[i] = std::atof(in.tok(0).c_str());
        r.eta[i] = std::atof(in.tok(1).c_str());
    }
}

void read_angular(Lines& in, AngularSpec& a) {
    in.next("BasisA");
    const std::string phi = in.tok(4);
    const int n = std::atoi(in.tok(5).c_str());
    a.alpha = std::atoi(in.tok(6).c_str());
    a.cut = make_cutoff(in.tok(1), in.tok(2), std::atof(in.tok(3).c_str()));
    if (phi == "GAUSS") a.phi = ANG_GAUSS;
    else if (phi == "IPOWP") a.phi = ANG_IPOWP;
    else if (phi == "IPOWS") a.phi = ANG_IPOWS;
    else if (phi == "SECHP") a.phi = ANG_SECHP;
    else if (phi == "SECHS") a.phi = ANG_SECHS;
    else  // basis_angular.cpp:62
        throw std::invalid_argument("BasisA(rc,cutoff::Name,cutoff::Norm,int,PhiAN): invalid angular function type");
    if (n < 0) throw std::invalid_argument("Basis::resize(int): invalid number of functions.");
    a.eta.resize(n);
    a.zeta.resize(n);
    a.lambda.resize(n);
    a.ietap.resize(n);
    for (int i = 0; i < n; ++i) {
        in.next("BasisA parameters");
        a.eta[i] = std::atof(in.tok(0).c_str());
        a.zeta[i] = std::atof(in.tok(1).c_str());
        a.lambda[i] = std::atoi(in.tok(2).c_str());
        a.ietap[i] = powint(1.0 / a.eta[i], a.alpha);  // basis_angular.cpp:171
    }
}

void read_net(Lines& in, NetSpec& net) {
    const auto& t = in.next("nn layout");
    std::vector<int> nodes;
    for (size_t i = 1; i < t.size(); ++i) nodes.push_back(std::atoi(t[i].c_str()));
    if (nodes.size() < 2) throw std::invalid_argument("ANN::resize(const ANNP&,int,const std::vector<int>&): Invalid number of layers.");
    for (int v : nodes)
        if (v <= 0) throw std::invalid_argument("ANN::resize(const ANNP&,int,const std::vector<int>&): Invalid layer size.");
    in.next("neuron");
    const std::string neuron = in.tok(1);
    static const char* names[] = {"LINEAR", "SIGMOID", "TANH", "ISRU", "ARCTAN", "SOFTSIGN", "RELU", "SOFTPLUS", "ELU",
                                  "GELU", "SWISH", "MISH", "TANHRE", "LOGCOSH", "IQUAD", "IFABS", "SINSIG"};
    net.neuron = 0;
    for (int k = 0; k < 17; ++k)
        if (neuron == names[k]) net.neuron = k + 1;  // nn.cpp:115-134
    if (net.neuron == 0) throw std::invalid_argument("ANN::read(FILE*,ANN&): Invalid neuron.");  // nn.cpp:1376
    net.nin = nodes[0];
    net.nodes.assign(nodes.begin() + 1, nodes.end());
    if (net.nodes.back() != 1) throw unsupported_error("a potential network must have exactly one output");
    if ((int)net.nodes.size() > kMaxLayers) throw unsupported_error("network deeper than the engine supports");
    for (int v : nodes)
        if (v > kMaxWidth) throw unsupported_error("network layer wider than the engine supports");
    auto vec = [&](const char* what, std::vector<double>& v, int n) {
        in.next(what);
        v.resize(n);
        for (int i = 0; i < n; ++i) v[i] = std::atof(in.tok(1 + i).c_str());
    };
    vec("inpw", net.inpw, net.nin);
    vec("outw", net.outw, 1);
    vec("inpb", net.inpb, net.nin);
    vec("outb", net.outb, 1);
    const int L = (int)net.nodes.size();
    net.b.resize(L);
    net.W.resize(L);
    for (int l = 0; l < L; ++l) vec("bias", net.b[l], net.nodes[l]);
    for (int l = 0; l < L; ++l) vec("weight", net.W[l], net.nodes[l] * (l == 0 ? net.nin : net.nodes[l - 1]));
}

}  // namespace

HostModel parse_ann(const char* path) {
    FILE* f = std::fopen(path, "r");
    if (!f) throw io_error(std::string("NNP::read(const char*,NNP&): Could not open nnp file: \"") + path + "\"");
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
    Lines in(f);
    HostModel m;
    in.next("header");
    in.next("ntypes");
    m.T = std::atoi(in.tok(1).c_str());
    if (m.T <= 0) throw std::invalid_argument("NNP::read(FILE*,NNP&): invalid number of types.");
    if (m.T > kMaxSpecies) throw unsupported_error("more species than the engine supports");
    m.sp.resize(m.T);
    for (int n = 0; n < m.T; ++n) {  // type.cpp:32-70
        const auto& t = in.next("type");
        SpeciesSpec& s = m.sp[n];
        for (size_t i = 0; i < t.size(); ++i) {
            if (t[i][0] == '#') break;
            std::string tag = t[i];
            for (auto& c : tag) c = (char)std::toupper((unsigned char)c);
            if (tag == "NAME") s.name = in.tok(++i);
            else if (tag == "ENERGY") s.evac = std::atof(in.tok(++i).c_str());
            else if (tag == "MASS" || tag == "CHARGE" || tag == "CHI" || tag == "ETA" || tag == "JS" || tag == "C6" ||
                     tag == "RVDW" || tag == "RCOV")
                ++i;
        }
        s.hash = name_hash(s.name.c_str());
    }
    in.next("rc");
    m.rc = std::atof(in.tok(1).c_str());
    if (m.rc <= 0) throw std::invalid_argument("NNP::read(FILE*,NNP&): invalid cutoff.");
    const int nblk = m.T * (m.T + 1) / 2;
    for (auto& s : m.sp) {
        s.radial.assign(m.T, RadialSpec());
        s.angular.assign(nblk, AngularSpec());
    }
    auto lookup = [&](const std::string& name) {
        const int k = m.index(name);
        if (k < 0) throw std::runtime_error("No match for key in map");  // mem/map.hpp:122
        return k;
    };
    for (int i = 0; i < m.T; ++i) {
        in.next("basis");
        const int II = lookup(in.tok(1));
        for (int j = 0; j < m.T; ++j) {
            in.next("basis_radial");
            read_radial(in, m.sp[II].radial[lookup(in.tok(1))]);
        }
        for (int j = 0; j < m.T; ++j)
            for (int k = j; k < m.T; ++k) {
                in.next("basis_angular");
                const int JJ = lookup(in.tok(1)), KK = lookup(in.tok(2));
                read_angular(in, m.sp[II].angular[lmat_index(JJ, KK)]);
            }
    }
    for (auto& s : m.sp) {  // nnh.cpp:95-121
        s.nR = 0;
        s.offR.assign(m.T, 0);
        for (int t = 0; t < m.T; ++t) {
            s.offR[t] = s.nR;
            s.nR += (int)s.radial[t].eta.size();
        }
        s.nA = 0;
        s.offA.assign(nblk, 0);
        for (int b = 0; b < nblk; ++b) {
            s.offA[b] = s.nA;
            s.nA += (int)s.angular[b].eta.size();
        }
        s.nG = s.nR + s.nA;
    }
    for (int n = 0; n < m.T; ++n) {
        in.next("nn");
        const int II = lookup(in.tok(1));
        read_net(in, m.sp[II].net);
        if (m.sp[II].net.nin != m.sp[II].nG)
            throw std::invalid_argument("network input size does not match the basis of species " + m.sp[II].name);
    }
    return m;
}

static int pad_to(int v, int q) { return (v + q - 1) / q * q; }

PlanBuffers build_plan(const HostModel& m) {
    PlanBuffers pb;
    DevPlan& P = pb.plan;
    std::memset(&P, 0, sizeof(P));
    P.T = m.T;
    P.ldg = 0;
    P.ne_max = 1;
    P.nc_max = 1;
    const int nblk = m.T * (m.T + 1) / 2;
    for (int i = 0; i < m.T; ++i) {
        const SpeciesSpec& s = m.sp[i];
        P.nG[i] = s.nG;
        P.nR[i] = s.nR;
        if (s.nG > P.ldg) P.ldg = s.nG;
        for (int j = 0; j < m.T; ++j) {
            const RadialSpec& r = s.radial[j];
            DevRadial& d = P.rad[i][j];
            d.n = (int)r.eta.size();
            d.goff = s.offR[j];
            d.fbase = (int)pb.rfun_eta.size();
            d.cut = r.cut.kind;
            d.phi = r.phi;
            if (r.phi != RAD_GAUSSIAN) P.generic = 1;
            d.rc = r.cut.rc;
            d.rci = r.cut.rci;
            d.pirci = r.cut.pirci;
            d.norm = r.cut.norm;
            for (int f = 0; f < d.n; ++f) {
                pb.rfun_eta.push_back(r.eta[f]);
                pb.rfun_rs.push_back(r.rs[f]);
            }
        }
        for (int b = 0; b < nblk; ++b) {
            const AngularSpec& a = s.angular[b];
            DevAngular& d = P.ang[i][b];
            d.goff = s.nR + s.offA[b];
            d.alpha = a.alpha;
            d.cut = a.cut.kind;
            d.phi = a.phi;
            if (a.phi != ANG_GAUSS) P.generic = 1;
            d.rc = a.cut.rc;
            d.rci = a.cut.rci;
            d.pirci = a.cut.pirci;
            d.norm = a.cut.norm;
            for (int q = 0; q < kMaxE * kMaxC; ++q) d.fmap[q] = -1;
            // factorise the function list into distinct exponents x distinct angular terms
            std::vector<double> E;
            std::vector<std::pair<double, int>> C;
            for (size_t f = 0; f < a.eta.size(); ++f) {
                int e = -1, c = -1;
                for (size_t q = 0; q < E.size(); ++q)
                    if (E[q] == a.ietap[f]) e = (int)q;
                if (e < 0) {
                    E.push_back(a.ietap[f]);
                    e = (int)E.size() - 1;
                }
                for (size_t q = 0; q < C.size(); ++q)
                    if (C[q].first == a.zeta[f] && C[q].second == a.lambda[f]) c = (int)q;
                if (c < 0) {
                    C.push_back({a.zeta[f], a.lambda[f]});
                    c = (int)C.size() - 1;
                }
                if (e >= kMaxE || c >= kMaxC)
                    throw unsupported_error("angular basis has more distinct exponents or (zeta,lambda) terms than the engine supports");
                if (d.fmap[e * kMaxC + c] >= 0)
                    throw unsupported_error("angular basis repeats an (eta, zeta, lambda) function");
                d.fmap[e * kMaxC + c] = (int)f;
            }
            d.ne = (int)E.size();
            d.nc = (int)C.size();
            for (int e = 0; e < kMaxE; ++e) d.ietap[e] = e < d.ne ? E[e] : 0.0;
            for (int c = 0; c < kMaxC; ++c) {
                d.zeta[c] = c < d.nc ? C[c].first : 1.0;
                d.lambda[c] = c < d.nc ? (double)C[c].second : 1.0;
                const double z = d.zeta[c];
                d.zint[c] = (z >= 0 && z <= 64 && std::floor(z) == z) ? (int)z : -1;
                // the fast kernels sum the angular terms of a block as one polynomial in cos of
                // degree <= 4: fractional zeta, zeta > 4 or |lambda| > 1 take the generic kernels
                if (c < d.nc && (d.zint[c] < 0 || d.zint[c] > 4 || std::fabs(d.lambda[c]) > 1.0)) P.generic = 1;
            }
            if (d.ne > P.ne_max) P.ne_max = d.ne;
            if (d.nc > P.nc_max) P.nc_max = d.nc;
        }
        // moment expansion eligibility of this centre species
        {
            bool shared = true, ok = true;
            const DevAngular& A0 = P.ang[i][0];
            for (int b = 0; b < nblk && shared; ++b) {
                const DevAngular& A = P.ang[i][b];
                if (A.ne != A0.ne || A.alpha != A0.alpha || A.cut != A0.cut || A.rc != A0.rc || A.phi != A0.phi ||
                    !angular_separable(A.phi))
                    shared = false;
                for (int e = 0; e < A.ne && shared; ++e)
                    if (A.ietap[e] != A0.ietap[e]) shared = false;
                for (int c = 0; c < A.nc; ++c)
                    if (A.zint[c] < 0 || A.zint[c] > 4 || std::fabs(A.lambda[c]) > 1.0) ok = false;
            }
            P.ang_shared[i] = shared ? 1 : 0;
            P.mom_ok[i] = (shared && ok) ? 1 : 0;
        }
        // network: weights padded (outputs to 16, inputs to 16) and stored row-major
        const NetSpec& net = s.net;
        DevNet& dn = P.net[i];
        dn.L = (int)net.nodes.size();
        dn.nin = net.nin;
        dn.nin_pad = pad_to(net.nin, 16);
        dn.act = net.neuron;
        if (net.neuron != NEU_TANH && net.neuron != NEU_LINEAR) P.generic = 1;
        dn.outw = net.outw[0];
        dn.outb = net.outb[0];
        dn.evac = s.evac;
        std::vector<double>& W = pb.weights;
        auto align = [&]() {
            while (W.size() % 32) W.push_back(0.0);
        };
        align();
        dn.inpw_off = (int64_t)W.size();
        for (int k = 0; k < dn.nin_pad; ++k) W.push_back(k < net.nin ? net.inpw[k] : 0.0);
        dn.inpb_off = (int64_t)W.size();
        for (int k = 0; k < dn.nin_pad; ++k) W.push_back(k < net.nin ? net.inpb[k] : 0.0);
        int in_real = net.nin, in_pad = dn.nin_pad;
        for (int l = 0; l < dn.L; ++l) {
            const int out_real = net.nodes[l], out_pad = pad_to(out_real, 16);
            dn.width[l] = out_real;
            dn.wpad[l] = out_pad;
            align();
            dn.woff[l] = (int64_t)W.size();
            for (int o = 0; o < out_pad; ++o)
                for (int k = 0; k < in_pad; ++k)
                    W.push_back((o < out_real && k < in_real) ? net.W[l][(size_t)o * in_real + k] : 0.0);
            align();
            dn.wtoff[l] = (int64_t)W.size();
            for (int k = 0; k < in_pad; ++k)
                for (int o = 0; o < out_pad; ++o)
                    W.push_back((o < out_real && k < in_real) ? net.W[l][(size_t)o * in_real + k] : 0.0);
            align();
            dn.boff[l] = (int64_t)W.size();
            for (int o = 0; o < out_pad; ++o) W.push_back(o < out_real ? net.b[l][o] : 0.0);
            in_real = out_real;
            in_pad = out_pad;
        }
        // contiguous copy of the layer matrices (of the output layer only its one real row) and
        // biases (one vector copy into LDS).  Matrix
        // element (o, k) sits at o in_pad + (k ^ (o & 15)): the 16 rows a forward MFMA operand load
        // reads at one k, and the 16 columns a backward one reads in one row, fall in 16 different
        // LDS banks (row-major, every row of a K-step would start in the same bank).
        align();
        dn.pack_off = (int64_t)W.size();
        for (int l = 0; l < dn.L; ++l) {
            dn.pw_off[l] = (int)((int64_t)W.size() - dn.pack_off);
            const int ip = l == 0 ? dn.nin_pad : dn.wpad[l - 1];
            const int rows = l == dn.L - 1 ? 1 : dn.wpad[l];  // the output layer's one real row
            const size_t base = W.size();
            W.resize(base + (size_t)rows * ip);
            for (int o = 0; o < rows; ++o)
                for (int k = 0; k < ip; ++k) W[base + (size_t)o * ip + (k ^ (o & 15))] = W[dn.woff[l] + (int64_t)o * ip + k];
        }
        for (int l = 0; l < dn.L; ++l) {
            dn.pb_off[l] = (int)((int64_t)W.size() - dn.pack_off);
            for (int o = 0; o < dn.wpad[l]; ++o) {
                const double v = W[dn.boff[l] + o];
                W.push_back(v);
            }
        }
        if (W.size() % 2) W.push_back(0.0);
        dn.pack_len = (int)((int64_t)W.size() - dn.pack_off);
    }
    if (pb.rfun_eta.empty()) {
        pb.rfun_eta.push_back(0.0);
        pb.rfun_rs.push_back(0.0);
    }
    return pb;
}

}  // namespace atomnn

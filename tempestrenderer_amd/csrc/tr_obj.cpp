/*
 * tr_obj.cpp - Wavefront OBJ/MTL files into a ray-tracing mesh blob.
 *
 * LoadObjFileStaticRTGeometry (src/mesh/obj-loader.cc:635-840) with its
 * lexers (src/mesh/obj-loader.l, src/mesh/obj-mtl-loader.l, restated here
 * as a longest-match scanner with the same token rules and priorities),
 * parsers (src/mesh/obj-loader-parser.cc:322-515, obj-mtl-parser.cc:
 * 148-247), driver (obj-loader-driver.cc:36-86, obj-loader-driver.hh:
 * 51-83, obj-mtl-loader-driver.hh:48-54) and the per-group interleaving
 * InterleaveInterm (obj-loader.cc:61-384) over InterleaveVertices
 * (src/utils/interleave-vertices.cc:62-115).
 *
 * The result is what the reference fills into RTMeshBlob: per group ("g"
 * or "o") one submesh whose vertices are the distinct (position,
 * [texcoord], normal) index tuples of its faces in first-use order,
 * interleaved as position, [texcoord], [tangent], [consistent-normal
 * coefficient], normal; faces are fans; normals missing from the file are
 * the normalised sums of the face normals at each position.  Materials are
 * RTMicrofacetMaterial records (Blinn-Phong; Specular and SpecularPower
 * only for illum 2; Emissive when all three Ke components are non-zero,
 * the reference's IsZero test, vector3.hh:108-111), after setup().
 *
 * Where the reference's behaviour is undefined this loader fails with an
 * error instead: indices outside the arrays, groups whose faces mix vertex
 * formats.  Material and file names after usemtl / mtllib / map_* are
 * taken without surrounding blanks and carriage returns (the reference
 * keeps them, so such files fail to find their materials there).  A
 * material's map_Kd image is read with tr_image_load from the OBJ file's
 * directory (obj-loader.cc:755-767) into the blob's textures; one that
 * does not load leaves the material without a map and is counted in
 * `ignored_maps`, as the reference logs it and goes on.  Like the
 * reference, any diagnostic of the OBJ and MTL grammars fails the whole
 * load; the first one is the error message.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "tr_host.h"

using namespace trk;

namespace {

struct V2 { float x, y; };
struct V3 { float x, y, z; };

inline V3 sub3(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add3(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale3(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross3(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
/* NormalizeSelf (vector3.hh:78-88): times 1/length, zero vectors kept. */
inline void normalize_self(V3* v)
{
    float l = sqrtf(v->x * v->x + v->y * v->y + v->z * v->z);
    if (l != 0.0f) {
        float r = 1.0f / l;
        v->x *= r;
        v->y *= r;
        v->z *= r;
    }
}
/* Normalize (vector3.hh:215-219): Vector3 / length, which multiplies by
 * the reciprocal (vector3.hh:185-190); zero vectors kept. */
inline V3 normalized(V3 v)
{
    float l = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (l == 0.0f) return v;
    float r = 1.0f / l;
    return V3{v.x * r, v.y * r, v.z * r};
}

/* ---------------- scanner ---------------- */

enum TokKind { TK_EOF, TK_EOL, TK_INT, TK_FLOAT, TK_STRING, TK_SLASH, TK_KEYWORD };

struct Keyword {
    const char* text;
    int id;
    bool name_follows; /* the rule ends in [ \t] and switches to the rest-of-line name state */
};

struct Token {
    TokKind kind = TK_EOF;
    int kw = -1;
    int ival = 0;
    float fval = 0.0f;
    std::string text;
};

class Scanner {
public:
    Scanner(const std::string& src, const Keyword* kws, size_t n_kws, bool illum_rule)
        : s_(src), kws_(kws), n_kws_(n_kws), illum_rule_(illum_rule) {}

    int line() const { return line_; }
    const std::string& error() const { return err_; }

    /* false on an invalid character (the lexers' "." rule) */
    bool next(Token* t)
    {
        t->kind = TK_EOF;
        t->kw = -1;
        t->text.clear();
        const size_t n = s_.size();
        if (name_state_) {
            if (p_ >= n) return true;
            size_t a = 0;
            while (p_ + a < n && s_[p_ + a] != '\n') ++a;
            size_t r = 0;
            while (p_ + r < n && s_[p_ + r] == '\r') ++r;
            size_t b = (p_ + r < n && s_[p_ + r] == '\n') ? r + 1 : 0;
            if (b > a) {
                p_ += b;
                ++line_;
                name_state_ = false;
                t->kind = TK_EOL;
                return true;
            }
            t->kind = TK_STRING;
            t->text = s_.substr(p_, a);
            p_ += a;
            return true;
        }
        for (;;) {
            while (p_ < n && (s_[p_] == ' ' || s_[p_] == '\t')) ++p_;
            if (p_ < n && s_[p_] == '#') {
                while (p_ < n && s_[p_] != '\r' && s_[p_] != '\n') ++p_;
                continue;
            }
            break;
        }
        if (p_ >= n) return true;
        const char c = s_[p_];
        if (c == '\r' || c == '\n') {
            size_t r = 0;
            while (p_ + r < n && s_[p_ + r] == '\r') ++r;
            if (p_ + r < n && s_[p_ + r] == '\n') {
                p_ += r + 1;
                ++line_;
                t->kind = TK_EOL;
                return true;
            }
            return invalid(c);
        }
        if (c == '/') {
            ++p_;
            t->kind = TK_SLASH;
            return true;
        }
        /* longest match; ties go to the earlier rule: numbers, keywords, strings */
        bool is_float = false;
        size_t num = number_length(&is_float);
        size_t kw_len = 0;
        int kw_id = -1;
        bool kw_name = false;
        for (size_t k = 0; k < n_kws_; ++k) {
            size_t l = strlen(kws_[k].text);
            if (s_.compare(p_, l, kws_[k].text) != 0) continue;
            if (kws_[k].name_follows) {
                if (p_ + l >= n || (s_[p_ + l] != ' ' && s_[p_ + l] != '\t')) continue;
                ++l;
            }
            if (l > kw_len) {
                kw_len = l;
                kw_id = kws_[k].id;
                kw_name = kws_[k].name_follows;
            }
        }
        int illum_code = -1;
        if (illum_rule_ && s_.compare(p_, 6, "illum_") == 0 && p_ + 6 < n && isdigit((unsigned char)s_[p_ + 6])) {
            if (7 > kw_len) {
                kw_len = 7;
                kw_id = -2;
                kw_name = false;
                illum_code = (unsigned char)s_[p_ + 6]; /* IntegerValue = yytext[6] */
            }
        }
        size_t str = 0;
        while (p_ + str < n && is_name_char(s_[p_ + str])) ++str;
        if (num == 0 && kw_len == 0 && str == 0) return invalid(c);
        if (num >= kw_len && num >= str) {
            std::string txt = s_.substr(p_, num);
            p_ += num;
            if (is_float) {
                t->kind = TK_FLOAT;
                t->fval = (float)strtod(txt.c_str(), nullptr);
            } else {
                t->kind = TK_INT;
                t->ival = (int)strtol(txt.c_str(), nullptr, 10);
            }
            t->text = txt;
            return true;
        }
        if (kw_len >= str) {
            t->text = s_.substr(p_, kw_len);
            p_ += kw_len;
            if (kw_id == -2) {
                t->kind = TK_INT;  /* the MTL IllumModel token carries an integer */
                t->kw = -2;
                t->ival = illum_code;
                return true;
            }
            t->kind = TK_KEYWORD;
            t->kw = kw_id;
            if (kw_name) name_state_ = true;
            return true;
        }
        t->kind = TK_STRING;
        t->text = s_.substr(p_, str);
        p_ += str;
        return true;
    }

private:
    static bool is_name_char(char c) { return isalnum((unsigned char)c) || c == '_' || c == '.'; }

    /* INTEGER "." [0-9]+ MANTISA? | INTEGER MANTISA | INTEGER, with
     * INTEGER = -?[0-9]+ and MANTISA = [eE][+-]?[0-9]* */
    size_t number_length(bool* is_float) const
    {
        const size_t n = s_.size();
        size_t q = p_;
        if (q < n && s_[q] == '-') ++q;
        size_t d0 = q;
        while (q < n && isdigit((unsigned char)s_[q])) ++q;
        if (q == d0) return 0;
        *is_float = false;
        if (q + 1 < n && s_[q] == '.' && isdigit((unsigned char)s_[q + 1])) {
            q += 1;
            while (q < n && isdigit((unsigned char)s_[q])) ++q;
            *is_float = true;
        }
        if (q < n && (s_[q] == 'e' || s_[q] == 'E')) {
            q += 1;
            if (q < n && (s_[q] == '+' || s_[q] == '-')) ++q;
            while (q < n && isdigit((unsigned char)s_[q])) ++q;
            *is_float = true;
        }
        return q - p_;
    }

    bool invalid(char c)
    {
        char buf[64];
        snprintf(buf, sizeof(buf), "invalid character: 0x%02x", (unsigned)(unsigned char)c);
        err_ = buf;
        return false;
    }

    const std::string& s_;
    const Keyword* kws_;
    size_t n_kws_;
    bool illum_rule_;
    size_t p_ = 0;
    int line_ = 1;
    bool name_state_ = false;
    std::string err_;
};

std::string trim_name(const std::string& s)
{
    size_t b = 0, e = s.size();
    while (b < e && (s[b] == ' ' || s[b] == '\t' || s[b] == '\r')) ++b;
    while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) --e;
    return s.substr(b, e - b);
}

std::string lower(std::string s)
{
    for (auto& c : s) c = (char)tolower((unsigned char)c);
    return s;
}

bool read_file(const std::string& path, std::string* out)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t r;
    out->clear();
    while ((r = fread(buf, 1, sizeof(buf), f)) > 0) out->append(buf, r);
    fclose(f);
    return true;
}

/* ---------------- MTL ---------------- */

/* ObjMtlLoader::Material (obj-mtl-loader.hh:57-76) with its defaults */
struct MtlMaterial {
    std::string name;
    V3 emission{0.0f, 0.0f, 0.0f};
    V3 ambient{1.0f, 1.0f, 1.0f};
    V3 diffuse{1.0f, 1.0f, 1.0f};
    V3 specular{0.0f, 0.0f, 0.0f};
    V3 transmission{0.0f, 0.0f, 0.0f};
    int illum = 0;
    float dissolve = 0.0f;
    float specular_exponent = 1.0f;
    float sharpness = 1.0f;
    float refraction_index = 1.0f;
    std::string map_ka, map_kd, map_ks, map_ns, map_d;
};

enum {
    M_NEWMTL, M_KA, M_KD, M_KS, M_TF, M_ILLUM, M_NS, M_SHARPNESS, M_NI, M_MAP_KA, M_MAP_KD, M_MAP_KS, M_MAP_NS,
    M_MAP_D, M_D, M_TR, M_KE
};
const Keyword kMtlKeywords[] = {
    {"newmtl", M_NEWMTL, false}, {"Ka", M_KA, false}, {"Kd", M_KD, false}, {"Ks", M_KS, false},
    {"Tf", M_TF, false}, {"illum", M_ILLUM, false}, {"Ns", M_NS, false}, {"sharpness", M_SHARPNESS, false},
    {"Ni", M_NI, false}, {"map_Ka", M_MAP_KA, true}, {"map_Kd", M_MAP_KD, true}, {"map_Ks", M_MAP_KS, true},
    {"map_Ns", M_MAP_NS, true}, {"map_d", M_MAP_D, true}, {"d", M_D, false}, {"Tr", M_TR, false},
    {"Ke", M_KE, false},
};

struct ParseError {
    std::string msg;
};

class TokenReader {
public:
    TokenReader(Scanner& sc, const std::string& file) : sc_(sc), file_(file) {}
    Token next()
    {
        Token t;
        if (!sc_.next(&t)) fail(sc_.error());
        return t;
    }
    [[noreturn]] void fail(const std::string& what) const
    {
        throw ParseError{file_ + ":" + std::to_string(sc_.line()) + ": " + what};
    }
    float number()
    {
        Token t = next();
        if (t.kind == TK_FLOAT) return t.fval;
        if (t.kind == TK_INT && t.kw != -2) return (float)t.ival;
        fail("expecting a number");
    }
    int index()
    {
        Token t = next();
        if (t.kind != TK_INT || t.kw == -2) fail("expecting an index");
        return t.ival;
    }
    void slash()
    {
        Token t = next();
        if (t.kind != TK_SLASH) fail("expecting \"/\"");
    }
    static bool is_end(const Token& t) { return t.kind == TK_EOF || t.kind == TK_EOL; }
    void end()
    {
        if (!is_end(next())) fail("expecting the end of the statement");
    }

private:
    Scanner& sc_;
    const std::string& file_;
};

/* ObjMtlLoader::Parser::parse (obj-mtl-parser.cc:148-247) */
void parse_mtl(const std::string& path, std::vector<MtlMaterial>* mats)
{
    std::string src;
    if (!read_file(path, &src)) throw ParseError{"cannot read material file: " + path};
    Scanner sc(src, kMtlKeywords, sizeof(kMtlKeywords) / sizeof(kMtlKeywords[0]), true);
    TokenReader rd(sc, path);
    MtlMaterial* cur = nullptr;
    auto need = [&]() {
        if (!cur) rd.fail("material field without newmtl");
        return cur;
    };
    for (;;) {
        Token t = rd.next();
        if (t.kind == TK_EOF) break;
        if (t.kind == TK_EOL) continue;
        if (t.kind != TK_KEYWORD) rd.fail("unexpected \"" + t.text + "\" at top level");
        switch (t.kw) {
        case M_NEWMTL: {
            Token n = rd.next();
            if (n.kind != TK_STRING) rd.fail("expecting a material name");
            mats->push_back(MtlMaterial());
            mats->back().name = lower(n.text);
            cur = &mats->back();
            break;
        }
        case M_KA: case M_KD: case M_KS: case M_KE: case M_TF: {
            float x = rd.number(), y = rd.number(), z = rd.number();
            rd.end();
            V3 v{x, y, z};
            MtlMaterial* m = need();
            if (t.kw == M_KA) m->ambient = v;
            else if (t.kw == M_KD) m->diffuse = v;
            else if (t.kw == M_KS) m->specular = v;
            else if (t.kw == M_KE) m->emission = v;
            else m->transmission = v;
            break;
        }
        case M_ILLUM: {
            Token v = rd.next();
            if (v.kind != TK_INT) rd.fail("invalid illumination model declaration");
            rd.end();
            need()->illum = v.ival;
            break;
        }
        case M_D: case M_TR: {
            float x = rd.number();
            Token e = rd.next();
            if (e.kind == TK_INT || e.kind == TK_FLOAT) e = rd.next();
            if (!TokenReader::is_end(e)) rd.fail("expecting the end of the statement");
            need()->dissolve = x;
            break;
        }
        case M_NS: case M_SHARPNESS: case M_NI: {
            float x = rd.number();
            rd.end();
            MtlMaterial* m = need();
            if (t.kw == M_NS) m->specular_exponent = x;
            else if (t.kw == M_SHARPNESS) m->sharpness = x;
            else m->refraction_index = x;
            break;
        }
        default: { /* map_* name */
            Token n = rd.next();
            if (n.kind != TK_STRING) rd.fail("expecting a file name");
            rd.end();
            MtlMaterial* m = need();
            std::string name = trim_name(n.text);
            if (t.kw == M_MAP_KA) m->map_ka = name;
            else if (t.kw == M_MAP_KD) m->map_kd = name;
            else if (t.kw == M_MAP_KS) m->map_ks = name;
            else if (t.kw == M_MAP_NS) m->map_ns = name;
            else m->map_d = name;
            break;
        }
        }
    }
}

/* ---------------- OBJ ---------------- */

enum {
    O_VP, O_VN, O_VT, O_V, O_DEG, O_BMAT, O_STEP, O_CSTYPE, O_PARM, O_TRIM, O_HOLE, O_SCRV, O_SP, O_END, O_CON,
    O_P, O_L, O_F, O_CURV, O_CURV2, O_SURF, O_G, O_S, O_MG, O_O, O_BEVEL, O_C_INTERP, O_D_INTERP, O_LOD,
    O_USEMTL, O_MTLLIB, O_SHADOW_OBJ, O_TRACE_OBJ, O_CTECH, O_STECH, O_OFF
};
const Keyword kObjKeywords[] = {
    {"vp", O_VP, false}, {"vn", O_VN, false}, {"vt", O_VT, false}, {"v", O_V, false}, {"deg", O_DEG, false},
    {"bmat", O_BMAT, false}, {"step", O_STEP, false}, {"cstype", O_CSTYPE, false}, {"parm", O_PARM, false},
    {"trim", O_TRIM, false}, {"hole", O_HOLE, false}, {"scrv", O_SCRV, false}, {"sp", O_SP, false},
    {"end", O_END, false}, {"con", O_CON, false}, {"p", O_P, false}, {"l", O_L, false}, {"f", O_F, false},
    {"curv", O_CURV, false}, {"curv2", O_CURV2, false}, {"surf", O_SURF, false}, {"g", O_G, false},
    {"s", O_S, false}, {"mg", O_MG, false}, {"o", O_O, false}, {"bevel", O_BEVEL, false},
    {"c_interp", O_C_INTERP, false}, {"d_interp", O_D_INTERP, false}, {"lod", O_LOD, false},
    {"usemtl", O_USEMTL, true}, {"mtllib", O_MTLLIB, true}, {"shadow_obj", O_SHADOW_OBJ, false},
    {"trace_obj", O_TRACE_OBJ, false}, {"ctech", O_CTECH, false}, {"stech", O_STECH, false}, {"off", O_OFF, false},
};

struct Group {
    uint32_t pos_start, tc_start, norm_start, material;
};

struct ObjData {
    std::vector<V3> pos;
    std::vector<V2> tc;
    std::vector<V3> norm;
    std::vector<int32_t> pos_ind, tc_ind, norm_ind;
    std::vector<Group> groups;
    std::vector<MtlMaterial> materials;
    uint32_t current_material = 0;
};

int32_t resolve(int32_t idx, size_t count) { return idx >= 0 ? idx - 1 : (int32_t)count + idx; }

/* ObjLoader::Parser::parseIndices (obj-loader-parser.cc:186-320): a fan
 * over the face's vertices in one of the four index forms. */
void parse_face(TokenReader& rd, ObjData* d)
{
    auto push = [&](int p, int t, int n, bool has_t, bool has_n) {
        d->pos_ind.push_back(resolve(p, d->pos.size()));
        if (has_t) d->tc_ind.push_back(resolve(t, d->tc.size()));
        if (has_n) d->norm_ind.push_back(resolve(n, d->norm.size()));
    };
    int sp = rd.index();
    Token t = rd.next();
    if (t.kind == TK_INT) { /* f v v v ... */
        int pp = t.ival;
        int cp = rd.index();
        for (;;) {
            push(sp, 0, 0, false, false);
            push(pp, 0, 0, false, false);
            push(cp, 0, 0, false, false);
            pp = cp;
            t = rd.next();
            if (t.kind != TK_INT) break;
            cp = t.ival;
        }
    } else if (t.kind == TK_SLASH) {
        t = rd.next();
        if (t.kind == TK_INT) {
            int st = t.ival;
            t = rd.next();
            if (t.kind == TK_INT) { /* f v/t ... */
                int pp = t.ival;
                rd.slash();
                int pt = rd.index();
                int cp = rd.index();
                for (;;) {
                    rd.slash();
                    int ct = rd.index();
                    push(sp, st, 0, true, false);
                    push(pp, pt, 0, true, false);
                    push(cp, ct, 0, true, false);
                    pp = cp;
                    pt = ct;
                    t = rd.next();
                    if (t.kind != TK_INT) break;
                    cp = t.ival;
                }
            } else if (t.kind == TK_SLASH) { /* f v/t/n ... */
                int sn = rd.index();
                int pp = rd.index();
                rd.slash();
                int pt = rd.index();
                rd.slash();
                int pn = rd.index();
                int cp = rd.index();
                for (;;) {
                    rd.slash();
                    int ct = rd.index();
                    rd.slash();
                    int cn = rd.index();
                    push(sp, st, sn, true, true);
                    push(pp, pt, pn, true, true);
                    push(cp, ct, cn, true, true);
                    pp = cp;
                    pt = ct;
                    pn = cn;
                    t = rd.next();
                    if (t.kind != TK_INT) break;
                    cp = t.ival;
                }
            } else {
                rd.fail("invalid face declaration");
            }
        } else if (t.kind == TK_SLASH) { /* f v//n ... */
            int sn = rd.index();
            int pp = rd.index();
            rd.slash();
            rd.slash();
            int pn = rd.index();
            int cp = rd.index();
            for (;;) {
                rd.slash();
                rd.slash();
                int cn = rd.index();
                push(sp, 0, sn, false, true);
                push(pp, 0, pn, false, true);
                push(cp, 0, cn, false, true);
                pp = cp;
                pn = cn;
                t = rd.next();
                if (t.kind != TK_INT) break;
                cp = t.ival;
            }
        } else {
            rd.fail("invalid face declaration");
        }
    } else {
        rd.fail("invalid face declaration");
    }
    if (!TokenReader::is_end(t)) rd.fail("invalid face declaration");
}

std::string directory_of(const std::string& path)
{
    size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

/* ObjLoader::Parser::parse (obj-loader-parser.cc:322-515) */
void parse_obj(const std::string& path, ObjData* d)
{
    std::string src;
    if (!read_file(path, &src)) throw ParseError{"cannot read object file: " + path};
    Scanner sc(src, kObjKeywords, sizeof(kObjKeywords) / sizeof(kObjKeywords[0]), false);
    TokenReader rd(sc, path);
    const std::string dir = directory_of(path);
    for (;;) {
        Token t = rd.next();
        if (t.kind == TK_EOF) break;
        if (t.kind == TK_EOL) continue;
        if (t.kind != TK_KEYWORD) rd.fail("unexpected \"" + t.text + "\" at top level");
        switch (t.kw) {
        case O_P: case O_L: case O_S: {
            Token e = t;
            do { e = rd.next(); } while (!TokenReader::is_end(e));
            break;
        }
        case O_V: {
            float x = rd.number(), y = rd.number(), z = rd.number(), w = 1.0f;
            Token e = rd.next();
            if (e.kind == TK_INT || e.kind == TK_FLOAT) w = e.kind == TK_INT ? (float)e.ival : e.fval;
            else if (!TokenReader::is_end(e)) rd.fail("vertex position: v <number> <number> <number> [<number>]");
            float f = 1.0f / w;
            d->pos.push_back(V3{x * f, y * f, z * f});
            break;
        }
        case O_VN: {
            float x = rd.number(), y = rd.number(), z = rd.number();
            rd.end();
            d->norm.push_back(V3{x, y, z});
            break;
        }
        case O_VT: {
            float x = rd.number(), y = rd.number();
            Token e = rd.next();
            if (e.kind == TK_INT || e.kind == TK_FLOAT) e = rd.next();
            if (!TokenReader::is_end(e)) rd.fail("texture coordinate: vt <number> <number> [<number>]");
            d->tc.push_back(V2{x, y});
            break;
        }
        case O_G: case O_O: {
            Token n = rd.next();
            if (n.kind != TK_STRING) rd.fail("mesh group: g <string>");
            d->groups.push_back(Group{(uint32_t)d->pos_ind.size(), (uint32_t)d->tc_ind.size(),
                                      (uint32_t)d->norm_ind.size(), d->current_material});
            break;
        }
        case O_F:
            parse_face(rd, d);
            break;
        case O_USEMTL: {
            Token n = rd.next();
            if (n.kind != TK_STRING) rd.fail("material name: usemtl <string>");
            std::string name = lower(trim_name(n.text));
            size_t i = 0;
            while (i < d->materials.size() && d->materials[i].name != name) ++i;
            if (i == d->materials.size()) rd.fail("unknown material: " + name);
            d->current_material = (uint32_t)i;
            if (!d->groups.empty()) d->groups.back().material = (uint32_t)i;
            break;
        }
        case O_MTLLIB: {
            Token n = rd.next();
            if (n.kind != TK_STRING) rd.fail("material library: mtllib <string>");
            parse_mtl(dir + "/" + trim_name(n.text), &d->materials);
            break;
        }
        default:
            rd.fail("unsupported statement \"" + t.text + "\"");
        }
    }
}

/* ---------------- interleaving ---------------- */

struct TupleHash {
    size_t operator()(const std::vector<int32_t>& k) const
    {
        size_t h = 1469598103934665603ull;
        for (int32_t v : k) h = (h ^ (uint32_t)v) * 1099511628211ull;
        return h;
    }
};

/* GenerateTangentSpace (math/triangle.hh:52-92), with Matrix2::inverse
 * (matrix2.hh:121-127) and transformRotationInverse (:113-114). */
void tangent_space(V3 p0, V3 p1, V3 p2, V2 t0, V2 t1, V2 t2, V3* tangent)
{
    V3 e0 = sub3(p1, p0), e1 = sub3(p2, p0);
    V2 c0{t1.x - t0.x, t1.y - t0.y}, c1{t2.x - t0.x, t2.y - t0.y};
    float det = c0.x * c1.y - c0.y * c1.x;
    V2 i0{c1.y / det, -c0.y / det}, i1{-c1.x / det, c0.x / det};
    if (!std::isfinite(i0.x)) {
        V3 n = cross3(e0, e1);
        if (c0.x != 0.0f) {
            *tangent = scale3(add3(e0, e1), c0.x >= 0.0f ? 1.0f : -1.0f);
        } else {
            V3 b = scale3(add3(e0, e1), c0.y >= 0.0f ? 1.0f : -1.0f);
            *tangent = cross3(b, n);
        }
        return;
    }
    auto tri = [&](float a, float b) { return V2{a * i0.x + b * i0.y, a * i1.x + b * i1.y}; };
    V2 x = tri(e0.x, e1.x), y = tri(e0.y, e1.y), z = tri(e0.z, e1.z);
    *tangent = V3{x.x, y.x, z.x};
}

/* InterleaveInterm (obj-loader.cc:61-384) for one group; appends the
 * group's vertices to `data` and its group-local indices to `inds`. */
void interleave_group(const ObjData& d, const Group& g, size_t pos_n, size_t tc_n, size_t norm_n, uint32_t flags,
                      std::vector<uint8_t>* data, std::vector<int32_t>* inds, uint32_t* stride, const std::string& file)
{
    *stride = 0;
    if (pos_n == 0) return;
    if (pos_n % 3 != 0 || (tc_n != 0 && tc_n != pos_n) || (norm_n != 0 && norm_n != pos_n))
        throw ParseError{file + ": a group mixes face vertex formats"};
    const int32_t* pi = d.pos_ind.data() + g.pos_start;
    const int32_t* ti = tc_n ? d.tc_ind.data() + g.tc_start : nullptr;
    const int32_t* ni = norm_n ? d.norm_ind.data() + g.norm_start : nullptr;
    for (size_t i = 0; i < pos_n; ++i) {
        if (pi[i] < 0 || (size_t)pi[i] >= d.pos.size()) throw ParseError{file + ": position index out of range"};
        if (ti && (ti[i] < 0 || (size_t)ti[i] >= d.tc.size())) throw ParseError{file + ": texcoord index out of range"};
        if (ni && (ni[i] < 0 || (size_t)ni[i] >= d.norm.size())) throw ParseError{file + ": normal index out of range"};
    }
    const bool gen_tan = (flags & TR_OBJ_GENERATE_TANGENTS) && tc_n != 0;
    const bool consistent = (flags & TR_OBJ_GENERATE_CONSISTENT_NORMALS) != 0;

    int32_t min_ind = 0x7FFFFFFF, max_ind = 0;
    std::vector<int32_t> gen_inds;
    if (gen_tan || norm_n == 0) {
        for (size_t i = 0; i < pos_n; ++i) {
            min_ind = std::min(min_ind, pi[i]);
            max_ind = std::max(max_ind, pi[i]);
        }
        gen_inds.resize(pos_n);
        for (size_t i = 0; i < pos_n; ++i) gen_inds[i] = pi[i] - min_ind;
    }
    std::vector<V3> gen_tans;
    if (gen_tan) {
        gen_tans.assign((size_t)(max_ind - min_ind + 1), V3{0.0f, 0.0f, 0.0f});
        for (size_t i = 0; i < pos_n; i += 3) {
            V3 tan;
            tangent_space(d.pos[pi[i]], d.pos[pi[i + 1]], d.pos[pi[i + 2]], d.tc[ti[i]], d.tc[ti[i + 1]],
                          d.tc[ti[i + 2]], &tan);
            normalize_self(&tan);
            for (int k = 0; k < 3; ++k) gen_tans[gen_inds[i + k]] = add3(gen_tans[gen_inds[i + k]], tan);
        }
        for (auto& t : gen_tans) normalize_self(&t);
    }
    const V3* norms = d.norm.data();
    const int32_t* norm_idx = ni;
    size_t norm_count = norm_n;
    std::vector<V3> gen_norms;
    if (norm_n == 0) {
        gen_norms.assign((size_t)(max_ind - min_ind + 1), V3{0.0f, 0.0f, 0.0f});
        for (size_t i = 0; i < pos_n; i += 3) {
            V3 prev = d.pos[pi[i]], cur = d.pos[pi[i + 1]], next = d.pos[pi[i + 2]];
            V3 n = cross3(sub3(next, cur), sub3(prev, cur));
            for (int k = 0; k < 3; ++k) gen_norms[gen_inds[i + k]] = add3(gen_norms[gen_inds[i + k]], n);
        }
        for (auto& n : gen_norms) normalize_self(&n);
        norms = gen_norms.data();
        norm_idx = gen_inds.data();
        norm_count = pos_n;
    }
    std::vector<float> gen_coefs;
    std::vector<int32_t> coef_inds;
    if (consistent) {
        int32_t nmin = 0x7FFFFFFF, nmax = 0;
        for (size_t i = 0; i < norm_count; ++i) {
            nmin = std::min(nmin, norm_idx[i]);
            nmax = std::max(nmax, norm_idx[i]);
        }
        gen_coefs.assign((size_t)(nmax - nmin + 1), 1.0f);
        coef_inds.resize(norm_count);
        for (size_t i = 0; i < norm_count; i += 3) {
            V3 a = norms[norm_idx[i]], b = norms[norm_idx[i + 1]], c = norms[norm_idx[i + 2]];
            float cab = dot3(a, b), ccb = dot3(c, b), cac = dot3(a, c);
            float lowest = std::min(cab, ccb);
            if (cac < lowest) lowest = cac;
            for (int k = 0; k < 3; ++k) {
                int32_t gi = coef_inds[i + k] = norm_idx[i + k] - nmin;
                if (lowest < gen_coefs[gi]) gen_coefs[gi] = lowest;
            }
        }
        const float w_scale = 0.03632f;
        for (auto& c : gen_coefs) {
            float omc = 1.0f - c;
            c = acosf(c) * (1.0f + w_scale * omc * omc);
        }
    }

    /* InterleaveVertices: distinct index tuples in first-use order */
    std::vector<const uint8_t*> arrays;
    std::vector<uint32_t> sizes;
    std::vector<const int32_t*> idx;
    arrays.push_back((const uint8_t*)d.pos.data()); sizes.push_back(12); idx.push_back(pi);
    if (tc_n) { arrays.push_back((const uint8_t*)d.tc.data()); sizes.push_back(8); idx.push_back(ti); }
    if (gen_tan) { arrays.push_back((const uint8_t*)gen_tans.data()); sizes.push_back(12); idx.push_back(gen_inds.data()); }
    if (consistent) { arrays.push_back((const uint8_t*)gen_coefs.data()); sizes.push_back(4); idx.push_back(coef_inds.data()); }
    arrays.push_back((const uint8_t*)norms); sizes.push_back(12); idx.push_back(norm_idx);
    for (uint32_t s : sizes) *stride += s;

    const size_t data_start = data->size();
    std::unordered_map<std::vector<int32_t>, int32_t, TupleHash> remap;
    remap.reserve(pos_n);
    std::vector<int32_t> key(arrays.size());
    int32_t next_index = 0;
    for (size_t i = 0; i < pos_n; ++i) {
        for (size_t j = 0; j < arrays.size(); ++j) key[j] = idx[j][i];
        auto it = remap.find(key);
        if (it != remap.end()) {
            inds->push_back(it->second);
            continue;
        }
        for (size_t j = 0; j < arrays.size(); ++j) {
            const uint8_t* src = arrays[j] + (size_t)sizes[j] * key[j];
            data->insert(data->end(), src, src + sizes[j]);
        }
        inds->push_back(next_index);
        remap.emplace(key, next_index);
        ++next_index;
    }
    if (gen_tan) {
        /* tangent orthogonalised against the vertex normal (obj-loader.cc:
         * 341-358), over this group's vertices */
        const uint32_t to = 20, no = *stride - 12;
        for (size_t o = data_start; o < data->size(); o += *stride) {
            V3 t, n;
            memcpy(&t, data->data() + o + to, 12);
            memcpy(&n, data->data() + o + no, 12);
            t = normalized(sub3(t, scale3(n, dot3(n, t))));
            memcpy(data->data() + o + to, &t, 12);
        }
    }
}

bool all_nonzero(V3 v) { return v.x != 0.0f && v.y != 0.0f && v.z != 0.0f; }

}  // namespace

extern "C" {

tr_status tr_load_obj_rt_geometry(const char* filename, uint32_t flags, tr_mesh_blob* out)
{
    if (!filename || !out) return fail(TR_ERROR_INVALID_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    if (flags & ~(uint32_t)(TR_OBJ_GENERATE_TANGENTS | TR_OBJ_GENERATE_CONSISTENT_NORMALS))
        return fail(TR_ERROR_INVALID_ARG, "unknown OBJ loader flags");
    ObjData d;
    std::vector<uint8_t> data;
    std::vector<int32_t> inds;
    std::vector<tr_submesh> subs;
    try {
        parse_obj(filename, &d);
        const size_t ng = d.groups.size();
        subs.resize(ng);
        size_t prev_ind = 0, prev_vert = 0;
        for (size_t i = 0; i < ng; ++i) {
            size_t pn, tn, nn;
            if (i + 1 == ng) {
                pn = d.pos_ind.size() - d.groups[i].pos_start;
                tn = d.tc_ind.size() - d.groups[i].tc_start;
                nn = d.norm_ind.size() - d.groups[i].norm_start;
            } else {
                pn = d.groups[i + 1].pos_start - d.groups[i].pos_start;
                tn = d.groups[i + 1].tc_start - d.groups[i].tc_start;
                nn = d.groups[i + 1].norm_start - d.groups[i].norm_start;
            }
            uint32_t stride = 0;
            interleave_group(d, d.groups[i], pn, tn, nn, flags, &data, &inds, &stride, filename);
            tr_submesh& sm = subs[i];
            sm.material = nullptr;
            sm.vertex_count = (uint32_t)(inds.size() - prev_ind);
            sm.base_index = (uint32_t)prev_ind;
            sm.vertex_offset = (uint32_t)prev_vert;
            sm.stride = stride;
            prev_ind = inds.size();
            prev_vert = data.size();
        }
    } catch (const ParseError& e) {
        return fail(TR_ERROR_INVALID_ARG, e.msg);
    } catch (const std::bad_alloc&) {
        return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory loading an OBJ file");
    }
    if (data.size() > 0xFFFFFFFFull) return fail(TR_ERROR_UNSUPPORTED, "OBJ vertex data beyond 4 GiB");

    const uint32_t nm = (uint32_t)d.materials.size();
    std::vector<tr_image> images;
    std::vector<uint32_t> image_of(nm, 0);
    const std::string dir = directory_of(filename);
    for (uint32_t i = 0; i < nm; ++i) {
        if (d.materials[i].map_kd.empty()) continue;
        tr_image img;
        if (tr_image_load((dir + "/" + d.materials[i].map_kd).c_str(), &img) == TR_OK) {
            images.push_back(img);
            image_of[i] = (uint32_t)images.size();
        } else {
            ++out->ignored_maps;
        }
    }
    out->textures = images.empty() ? nullptr : (tr_image*)calloc(images.size(), sizeof(tr_image));
    if (!images.empty() && !out->textures) {
        for (auto& im : images) tr_image_free(&im);
        return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory for the mesh blob");
    }
    for (size_t i = 0; i < images.size(); ++i) out->textures[i] = images[i];
    out->texture_count = (uint32_t)images.size();
    out->materials = nm ? (tr_material*)calloc(nm, sizeof(tr_material)) : nullptr;
    out->submeshes = subs.empty() ? nullptr : (tr_submesh*)calloc(subs.size(), sizeof(tr_submesh));
    out->vertex_data = data.empty() ? nullptr : (uint8_t*)malloc(data.size());
    out->index_data = inds.empty() ? nullptr : (int32_t*)malloc(inds.size() * sizeof(int32_t));
    if ((nm && !out->materials) || (!subs.empty() && !out->submeshes) || (!data.empty() && !out->vertex_data) ||
        (!inds.empty() && !out->index_data)) {
        tr_mesh_blob_free(out);
        return fail(TR_ERROR_OUT_OF_MEMORY, "out of memory for the mesh blob");
    }
    for (uint32_t i = 0; i < nm; ++i) {
        const MtlMaterial& in = d.materials[i];
        tr_material* m = &out->materials[i];
        tr_material_init(m, TR_MODEL_BLINN_PHONG);
        const float kd[3] = {in.diffuse.x, in.diffuse.y, in.diffuse.z};
        memcpy(m->diffuse, kd, sizeof(kd));
        if (in.illum == 2) { /* SpecularDiffuseAmbient */
            const float ks[3] = {in.specular.x, in.specular.y, in.specular.z};
            memcpy(m->specular, ks, sizeof(ks));
            m->specular_power[0] = m->specular_power[1] = in.specular_exponent;
        } else {
            m->specular[0] = m->specular[1] = m->specular[2] = 0.0f;
            m->specular_power[0] = m->specular_power[1] = 1.0f;
        }
        m->diffuse_map = image_of[i];
        if (all_nonzero(in.emission)) {
            m->model = TR_MODEL_EMISSIVE;
            const float ke[3] = {in.emission.x, in.emission.y, in.emission.z};
            memcpy(m->diffuse, ke, sizeof(ke));
        }
        tr_material_setup(m);
    }
    out->material_count = nm;
    for (size_t i = 0; i < subs.size(); ++i) {
        out->submeshes[i] = subs[i];
        const uint32_t mi = d.groups[i].material;
        out->submeshes[i].material = mi < nm ? &out->materials[mi] : nullptr;
    }
    out->submesh_count = (uint32_t)subs.size();
    if (!data.empty()) memcpy(out->vertex_data, data.data(), data.size());
    if (!inds.empty()) memcpy(out->index_data, inds.data(), inds.size() * sizeof(int32_t));
    out->vertex_bytes = data.size();
    out->index_count = inds.size();
    out->stride = 0;
    return TR_OK;
}

void tr_mesh_blob_free(tr_mesh_blob* blob)
{
    if (!blob) return;
    free(blob->submeshes);
    free(blob->vertex_data);
    free(blob->index_data);
    free(blob->materials);
    for (uint32_t i = 0; i < blob->texture_count; ++i) tr_image_free(&blob->textures[i]);
    free(blob->textures);
    memset(blob, 0, sizeof(*blob));
}

}  // extern "C"

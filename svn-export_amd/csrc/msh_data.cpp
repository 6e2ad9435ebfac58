// $ElementData reader and MSH writer.  See msh_data.hh for the contract.
#include "msh_data.hh"

#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "common.hh"

namespace fl {

std::string element_id_map(const HostMesh &m, bool boundary, std::vector<int> &ids) {
    ids.clear();
    if (!boundary) {
        ids.reserve(m.el_id.size());
        int last = -1;
        for (int id : m.el_id) {
            if (last >= id) return "Element IDs in non-increasing order, ID: " + std::to_string(id);
            ids.push_back(last = id);
        }
    } else {
        ids.reserve(m.bc_id.size());
        int last = -1;
        for (int b = 0; b < m.n_bc_elements(); b++) {
            int id = m.bc_id[b];
            // implicit boundary elements get -1: no data row can match them
            if (b >= m.n_file_bc || id < 0) {
                ids.push_back(last = -1);
                continue;
            }
            if (last >= id) return "Element IDs in non-increasing order, ID: " + std::to_string(id);
            ids.push_back(last = id);
        }
    }
    return "";
}

// ---- tokenizer (tokenizer.cc:20-120) ------------------------------------------

std::string MshTokenizer::open(const char *path) {
    path_ = path;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return std::string("Can not open input file '") + path + "'.";
    std::fseek(fp, 0, SEEK_END);
    long len = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    buf_.assign(len > 0 ? (size_t)len : 0, '\0');
    size_t got = len > 0 ? std::fread(&buf_[0], 1, (size_t)len, fp) : 0;
    std::fclose(fp);
    buf_.resize(got);
    pos_ = 0;
    eof_ = fail_ = false;
    line_no_ = 0;
    line_.clear();
    toks_.clear();
    tok_ = 0;
    return "";
}

// Tokens of one line: blanks and tabs separate, '"' toggles quoting, '\' escapes
// the next '"', '\' or 'n'; empty tokens are dropped (Tokenizer::operator++).
void MshTokenizer::tokenize() {
    toks_.clear();
    tok_ = 0;
    bad_escape_ = false;
    std::string t;
    bool quoted = false;
    for (size_t i = 0; i < line_.size(); i++) {
        char c = line_[i];
        if (c == '\\') {
            char d = i + 1 < line_.size() ? line_[++i] : '\0';
            if (d == 'n')
                t += '\n';
            else if (d == '\\' || d == '"')
                t += d;
            else {
                bad_escape_ = true;  // reported when a caller reaches this token
                return;
            }
        } else if (!quoted && (c == ' ' || c == '\t')) {
            if (!t.empty()) toks_.push_back(t);
            t.clear();
        } else if (c == '"') {
            quoted = !quoted;
        } else {
            t += c;
        }
    }
    if (!t.empty()) toks_.push_back(t);
}

const std::string &MshTokenizer::cur() {
    if (eol() && bad_escape_)
        throw Error(-1, "Unknown escape sequence, file: '" + path_ + "', line: '" + std::to_string(line_no_) + "'.");
    if (eol())
        throw Error(-1, "Missing token, file: '" + path_ + "', line: '" + std::to_string(line_no_) + "', position: '" +
                            std::to_string(tok_) + "'.");
    return toks_[tok_];
}

bool MshTokenizer::next_line(bool warn_remaining) {
    if (warn_remaining && !eol() && warnings)
        warnings->push_back("Remaining tokens, file '" + path_ + "', line '" + std::to_string(line_no_) +
                            "', after token #" + std::to_string(tok_) + ".");
    if (eof_) return false;
    line_.clear();
    while (!eof_ && line_.empty()) {
        // std::getline: a missing final newline sets eof, reading past the end sets eof and fail
        if (pos_ >= buf_.size()) {
            eof_ = fail_ = true;
        } else {
            size_t nl = buf_.find('\n', pos_);
            if (nl == std::string::npos) {
                line_.assign(buf_, pos_, std::string::npos);
                pos_ = buf_.size();
                eof_ = true;
            } else {
                line_.assign(buf_, pos_, nl - pos_);
                pos_ = nl + 1;
            }
        }
        line_no_++;
        size_t a = 0, b = line_.size();
        while (a < b && std::isspace((unsigned char)line_[a])) a++;
        while (b > a && std::isspace((unsigned char)line_[b - 1])) b--;
        line_ = line_.substr(a, b - a);
        if (!line_.empty() && line_[0] == '#') line_.clear();  // comment pattern "#"
    }
    if (!fail_) {
        tokenize();
        return true;
    }
    toks_.clear();
    tok_ = 0;
    return false;
}

bool MshTokenizer::skip_to(const char *pattern) {
    while (!eof_) {
        if (line_.find(pattern) != std::string::npos) {
            tokenize();
            return true;
        }
        next_line(false);
    }
    return false;
}

// boost::lexical_cast<unsigned int> / <double>: the whole token must convert.
static bool to_uint(const std::string &s, unsigned &v) {
    size_t i = s[0] == '+' ? 1 : 0;
    if (i >= s.size()) return false;
    unsigned long long r = 0;
    for (; i < s.size(); i++) {
        if (s[i] < '0' || s[i] > '9') return false;
        r = r * 10 + (unsigned)(s[i] - '0');
        if (r > 0xffffffffull) return false;
    }
    v = (unsigned)r;
    return true;
}
static bool to_double(const std::string &s, double &v) {
    size_t k = (s[0] == '+' || s[0] == '-') ? 1 : 0;
    if (s.size() > k + 1 && s[k] == '0' && (s[k + 1] == 'x' || s[k + 1] == 'X')) return false;
    char *e = nullptr;
    v = std::strtod(s.c_str(), &e);
    return e && *e == '\0' && e != s.c_str();
}

// ---- reader (msh_gmshreader.cc:250-397) ------------------------------------------

std::string ElementDataReader::open(const char *path) {
    std::string err = tok_.open(path);
    tok_.warnings = &warnings;
    last_ = MshDataHeader();
    last_.time = -std::numeric_limits<double>::infinity();
    last_.actual = false;
    opened_ = err.empty();
    return err;
}

std::string ElementDataReader::read_header(MshDataHeader &h) {
    auto bad = [&]() {
        return "Wrong format of the $ElementData header, file '" + tok_.file() + "', line " +
               std::to_string(tok_.line_num()) + ".";
    };
    unsigned n;
    tok_.next_line(false);
    if (!to_uint(tok_.cur(), n)) return bad();
    tok_.advance();
    h.field_name.clear();
    h.interpolation.clear();
    if (n > 0) {
        tok_.next_line(true), n--;
        h.field_name = tok_.cur();
        tok_.advance();
    }
    if (n > 0) {
        tok_.next_line(true), n--;
        h.interpolation = tok_.cur();
        tok_.advance();
    }
    for (; n > 0; n--) tok_.next_line(false);

    tok_.next_line(true);
    if (!to_uint(tok_.cur(), n)) return bad();
    tok_.advance();
    h.time = 0.0;
    if (n > 0) {
        tok_.next_line(true), n--;
        if (!to_double(tok_.cur(), h.time)) return bad();
        tok_.advance();
    }
    for (; n > 0; n--) tok_.next_line(false);

    tok_.next_line(true);
    if (!to_uint(tok_.cur(), n)) return bad();
    tok_.advance();
    h.time_index = 0;
    h.n_components = 1;
    h.n_entities = 0;
    unsigned *dst[3] = {&h.time_index, &h.n_components, &h.n_entities};
    for (int k = 0; k < 3 && n > 0; k++) {
        tok_.next_line(true), n--;
        if (!to_uint(tok_.cur(), *dst[k])) return bad();
        tok_.advance();
    }
    for (; n > 0; n--) tok_.next_line(false);
    return "";
}

std::string ElementDataReader::read(MshDataHeader &search, double *data, const std::vector<int> &ids) {
    if (!opened_) return "element data reader is not open";
    warnings.clear();
    char msg[512];
    try {
        while (last_.time <= search.time * (1.0 + 2.0 * DBL_EPSILON)) {
            if (last_.actual) {
                // the pending section's time has passed: apply its rows
                n_read = 0;
                size_t it = 0;
                unsigned i_row;
                for (i_row = 0; i_row < last_.n_entities; ++i_row) {
                    tok_.next_line(true);
                    unsigned id;
                    if (!to_uint(tok_.cur(), id))
                        return "Wrong format of $ElementData line, file '" + tok_.file() + "', line " +
                               std::to_string(tok_.line_num()) + ".";
                    tok_.advance();
                    while (it < ids.size() && ids[it] < (int)id) ++it;  // ids missing in the file keep their values
                    if (it == ids.size()) {
                        std::snprintf(msg, sizeof msg,
                                      "In file '%s', '$ElementData' section for field '%s', time: %f. Data ID %u not "
                                      "found or is not in order. Skipping rest of data.",
                                      tok_.file().c_str(), search.field_name.c_str(), last_.time, id);
                        warnings.push_back(msg);
                        break;
                    }
                    if (ids[it] == (int)id) {
                        double *dp = data + it * (size_t)search.n_components;
                        for (unsigned c = 0; c < search.n_components; ++c, ++dp) {
                            if (!to_double(tok_.cur(), *dp))
                                return "Wrong format of $ElementData line, file '" + tok_.file() + "', line " +
                                       std::to_string(tok_.line_num()) + ".";
                            tok_.advance();
                        }
                        n_read++;
                    }
                    // a row whose id is not in the mesh is skipped
                }
                // after an early end the loop below also passes the line that ended it
                while (i_row < last_.n_entities) tok_.next_line(false), ++i_row;
                search.actual = true;
            }

            // next section of the same field
            last_.field_name.clear();
            while (!tok_.eof() && last_.field_name != search.field_name) {
                if (tok_.skip_to("$ElementData")) {
                    std::string err = read_header(last_);
                    if (!err.empty()) return err;
                }
            }
            if (tok_.eof()) {
                if (!last_.actual)
                    return "In file '" + tok_.file() + "', missing '$ElementData' section for field '" +
                           search.field_name + "'.";
                last_.time = std::numeric_limits<double>::infinity();  // the last section holds for all later times
                return "";
            }
            if (last_.n_components != search.n_components) {
                std::snprintf(msg, sizeof msg,
                              "In file '%s', '$ElementData' section for field '%s', time: %f. Wrong number of "
                              "components: %u, using %u instead.",
                              tok_.file().c_str(), search.field_name.c_str(), last_.time, last_.n_components,
                              search.n_components);
                warnings.push_back(msg);
                last_.n_components = search.n_components;
            }
            if (last_.n_entities != search.n_entities) {
                std::snprintf(msg, sizeof msg,
                              "In file '%s', '$ElementData' section for field '%s', time: %f. Wrong number of "
                              "entities: %u, using %u instead.",
                              tok_.file().c_str(), search.field_name.c_str(), last_.time, last_.n_entities,
                              search.n_entities);
                warnings.push_back(msg);
            }
            last_.actual = true;
        }
    } catch (const Error &e) {
        return e.what();
    }
    return "";
}

// ---- writer (output_msh.cc:53-98, 100-188, 383-408) ---------------------------------

std::string msh_write_head(const HostMesh &m, const char *path) {
    FILE *f = std::fopen(path, "w");
    if (!f) return std::string("can not open output file ") + path;
    static const int kGmshType[4] = {0, 1, 2, 4};
    std::fprintf(f, "$MeshFormat\n2 0 %d\n$EndMeshFormat\n", (int)sizeof(double));
    std::fprintf(f, "$Nodes\n%d\n", m.n_nodes());
    for (int i = 0; i < m.n_nodes(); i++)
        std::fprintf(f, "%d %g %g %g\n", i + 1, m.xyz[3 * i], m.xyz[3 * i + 1], m.xyz[3 * i + 2]);
    std::fprintf(f, "$EndNodes\n$Elements\n%d\n", m.n_elements());
    for (int e = 0; e < m.n_elements(); e++) {
        int rid = m.regions[m.el_region[e]].id, d = m.el_dim[e];
        std::fprintf(f, "%d %d 3 %d %d %d", e + 1, kGmshType[d], rid, rid, m.el_pid.empty() ? 0 : m.el_pid[e]);
        for (int k = 0; k <= d; k++) std::fprintf(f, " %d", m.el_nodes[4 * e + k] + 1);
        std::fputc('\n', f);
    }
    std::fprintf(f, "$EndElements\n");
    bool ok = !std::ferror(f);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? "" : std::string("write error on ") + path;
}

static std::string write_data_section(const char *section, const char *path, const char *name, const char *units,
                                      double time, int step, int n_comp, int n_values, const double *values,
                                      int precision) {
    if (n_comp < 1 || n_values < 0) return "bad data shape";
    if (precision < 1 || precision > 17) return "precision must be in 1..17";
    FILE *f = std::fopen(path, "a");
    if (!f) return std::string("can not open output file ") + path;
    const int p = precision;
    std::fprintf(f, "$%s\n1\n\"%s_[%s]\"\n1\n%.*g\n3\n%d\n%d\n%d\n", section, name, units, p, time, step, n_comp,
                 n_values);
    for (int i = 0; i < n_values; i++) {
        const double *v = values + (size_t)i * n_comp;
        if (n_comp == 1) {
            std::fprintf(f, "%d %.*g\n", i + 1, p, v[0]);
        } else {
            // vector rows end with a blank before the newline
            std::fprintf(f, "%d ", i + 1);
            for (int c = 0; c < n_comp; c++) std::fprintf(f, "%.*g ", p, v[c]);
            std::fputc('\n', f);
        }
    }
    std::fprintf(f, "$End%s\n", section);
    bool ok = !std::ferror(f);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? "" : std::string("write error on ") + path;
}

std::string msh_write_element_data(const char *path, const char *name, const char *units, double time, int step,
                                   int n_comp, int n_values, const double *values, int precision) {
    return write_data_section("ElementData", path, name, units, time, step, n_comp, n_values, values, precision);
}

std::string msh_write_node_data(const char *path, const char *name, const char *units, double time, int step,
                                int n_comp, int n_values, const double *values, int precision) {
    return write_data_section("NodeData", path, name, units, time, step, n_comp, n_values, values, precision);
}

}  // namespace fl

// Element-wise data in gmsh 2.x ASCII files: the `$ElementData` reader that
// feeds per-element input fields (conductivity, initial concentration,
// sources) and the MSH writer of the per-element results.
//
// Reader contract (GmshMeshReader::read_element_data, msh_gmshreader.cc:306-397,
// driven by FieldElementwise::set_time, field_elementwise_impl.hh:92-117):
//  * the file is a stream of lines; blank lines and lines starting with '#'
//    are skipped, tokens are separated by blanks or tabs, '"' quotes a token
//    (tokenizer.cc:20-120);
//  * sections of the requested field are applied in file order while their
//    time tag is <= t * (1 + 2 eps); the reader keeps its place between calls,
//    so a later time continues from the first section not yet applied;
//  * data rows `id v1 .. vn` are matched against the mesh's element ids in
//    increasing order: rows whose id is not in the mesh are skipped, a row
//    whose id is beyond the last mesh id ends the section (warning);
//  * values stay in the caller's buffer between calls (zero at first use,
//    field_elementwise_impl.hh:121-137).
// Writer contract (OutputMSH, output_msh.cc:53-98, 100-188, 383-408): the mesh
// head, then one `$ElementData` section per output field and time, values
// with the stream's default format (%g, 6 significant digits) unless a
// precision is given.
#pragma once
#include <string>
#include <vector>

#include "mesh.hh"

namespace fl {

// Mesh::elements_id_maps (mesh.cc:781-812): gmsh ids of the bulk elements
// (strictly increasing, else an error), or of the boundary elements with -1
// for the implicit ones.  Returns "" or an error message.
std::string element_id_map(const HostMesh &m, bool boundary, std::vector<int> &ids);

struct MshDataHeader {
    std::string field_name, interpolation;
    double time = 0.0;
    unsigned time_index = 0, n_components = 1, n_entities = 0;
    bool actual = false;
};

class MshTokenizer {
   public:
    std::string open(const char *path);
    bool next_line(bool warn_remaining);
    bool skip_to(const char *pattern);
    bool eof() const { return eof_; }
    bool eol() const { return tok_ >= toks_.size(); }
    const std::string &cur();  // throws fl::Error when the line has no token left
    void advance() {
        if (!eol()) tok_++;
    }
    int line_num() const { return line_no_; }
    const std::string &file() const { return path_; }
    std::vector<std::string> *warnings = nullptr;

   private:
    void tokenize();
    std::string buf_, path_, line_;
    size_t pos_ = 0;
    bool eof_ = false, fail_ = false, bad_escape_ = false;
    int line_no_ = 0;
    std::vector<std::string> toks_;
    size_t tok_ = 0;
};

class ElementDataReader {
   public:
    std::string open(const char *path);
    // One FieldElementwise::set_time: reads into data[ids.size() * search.n_components];
    // search.actual reports whether data changed.  Returns "" or an error message.
    std::string read(MshDataHeader &search, double *data, const std::vector<int> &ids);
    std::vector<std::string> warnings;  // of the last read
    int n_read = 0;                     // rows stored by the last section applied

   private:
    std::string read_header(MshDataHeader &h);
    MshTokenizer tok_;
    MshDataHeader last_;
    bool opened_ = false;
};

// OutputMSH::write_head (output_msh.cc:53-98, 424-435): creates `path` with the
// $MeshFormat, $Nodes and $Elements of the bulk mesh.
std::string msh_write_head(const HostMesh &m, const char *path);

// OutputMSH::write_msh_elem_data (output_msh.cc:383-408) with the value rows of
// write_msh_ascii_cont_data (:100-188): appends one $ElementData section named
// "<name>_[<units>]", values[n_values * n_comp] row by row with ids 1..n_values.
std::string msh_write_element_data(const char *path, const char *name, const char *units, double time, int step,
                                   int n_comp, int n_values, const double *values, int precision);

// OutputMSH::write_msh_node_data (output_msh.cc:327-381): one $NodeData section,
// rows "id v1 .. vn" with ids 1..n_values, as the element data.
std::string msh_write_node_data(const char *path, const char *name, const char *units, double time, int step,
                                int n_comp, int n_values, const double *values, int precision);

}  // namespace fl

"""CPU restatement of the gmsh `$ElementData` reading rules, pure Python.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
reader (fl_element_data_*), never by the product path.

Follows, as plain data-flow over a list of lines:
  * Tokenizer (src/system/tokenizer.cc:20-120): lines are trimmed, blank lines
    and lines starting with '#' are skipped, tokens are split at blanks and
    tabs with '"' quoting, empty tokens are dropped;
  * GmshMeshReader::read_data_header (src/mesh/msh_gmshreader.cc:250-300):
    string tags (field name, interpolation), real tags (time), integer tags
    (time index, components, entities), surplus tags skipped;
  * GmshMeshReader::read_element_data (msh_gmshreader.cc:306-397) driven by
    FieldElementwise::set_time (src/fields/field_elementwise_impl.hh:92-117):
    sections of the field are applied in file order while their time is
    <= t (1 + 2 eps), rows matched against increasing element ids.
Warnings are collected as short tags ("entities", "components", "skip-rest")
so tests can compare them with the library's messages.
"""
import math
import sys

_EPS = sys.float_info.epsilon


def _tokens(line):
    out, t, q, i = [], "", False, 0
    while i < len(line):
        c = line[i]
        if c == "\\":
            i += 1
            d = line[i] if i < len(line) else ""
            if d == "n":
                t += "\n"
            elif d in ("\\", '"'):
                t += d
            else:
                raise ValueError("escape")
        elif not q and c in " \t":
            if t:
                out.append(t)
            t = ""
        elif c == '"':
            q = not q
        else:
            t += c
        i += 1
    if t:
        out.append(t)
    return out


class _Lines:
    """The reader's view of the file: a cursor over non-blank, non-comment lines.
    `eof` turns true once a read finds no further line (or the final line has no
    newline), as std::getline reports it."""

    def __init__(self, text):
        raw = text.split("\n")
        self.trailing_nl = text.endswith("\n")
        if self.trailing_nl:
            raw = raw[:-1]
        self.raw = raw
        self.k = 0
        self.eof = False
        self.line = ""
        self.toks = []
        self.t = 0

    def next_line(self):
        if self.eof:
            return False
        self.line = ""
        failed = False
        while not self.eof and self.line == "":
            if self.k >= len(self.raw):
                self.eof = failed = True
            else:
                s = self.raw[self.k]
                self.k += 1
                if self.k == len(self.raw) and not self.trailing_nl:
                    self.eof = True
                s = s.strip()
                self.line = "" if s.startswith("#") else s
        if failed:
            self.toks, self.t = [], 0
            return False
        self.toks, self.t = _tokens(self.line), 0
        return True

    def tok(self):
        if self.t >= len(self.toks):
            raise ValueError("missing token")
        v = self.toks[self.t]
        self.t += 1
        return v

    def skip_to(self, pattern):
        while not self.eof:
            if pattern in self.line:
                self.toks, self.t = _tokens(self.line), 0
                return True
            self.next_line()
        return False


def _uint(s):
    if not s.lstrip("+").isdigit() or s.count("+") > 1 or (s.startswith("+") and len(s) == 1):
        raise ValueError("uint")
    return int(s)


def _real(s):
    if "x" in s.lower() or "_" in s:
        raise ValueError("real")
    return float(s)


class ElementDataReader:
    def __init__(self, path):
        with open(path) as f:
            self.L = _Lines(f.read())
        self.last = dict(name="", time=-math.inf, ncomp=1, nent=0, actual=False)
        self.warnings = []
        self.n_read = 0

    def _header(self):
        L, h = self.L, self.last
        L.next_line()
        n = _uint(L.tok())
        h["name"] = ""
        if n > 0:
            L.next_line()
            n -= 1
            h["name"] = L.tok()
        if n > 0:
            L.next_line()
            n -= 1
            L.tok()  # interpolation scheme
        for _ in range(n):
            L.next_line()
        L.next_line()
        n = _uint(L.tok())
        h["time"] = 0.0
        if n > 0:
            L.next_line()
            n -= 1
            h["time"] = _real(L.tok())
        for _ in range(n):
            L.next_line()
        L.next_line()
        n = _uint(L.tok())
        vals = [0, 1, 0]
        for k in range(3):
            if n > 0:
                L.next_line()
                n -= 1
                vals[k] = _uint(L.tok())
        for _ in range(n):
            L.next_line()
        h["ncomp"], h["nent"] = vals[1], vals[2]

    def read(self, name, time, ncomp, ids, data):
        """One set_time: updates data[len(ids) * ncomp] in place, returns True when
        a section was applied."""
        L, h = self.L, self.last
        self.warnings = []
        changed = False
        while h["time"] <= time * (1.0 + 2.0 * _EPS):
            if h["actual"]:
                self.n_read = 0
                j = 0
                row = 0
                while row < h["nent"]:
                    L.next_line()
                    eid = _uint(L.tok())
                    while j < len(ids) and ids[j] < eid:
                        j += 1
                    if j == len(ids):
                        self.warnings.append("skip-rest")
                        break
                    if ids[j] == eid:
                        for c in range(ncomp):
                            data[j * ncomp + c] = _real(L.tok())
                        self.n_read += 1
                    row += 1
                while row < h["nent"]:
                    L.next_line()
                    row += 1
                changed = True
            h["name"] = ""
            while not L.eof and h["name"] != name:
                if L.skip_to("$ElementData"):
                    self._header()
            if L.eof:
                if not h["actual"]:
                    raise ValueError("missing $ElementData section for field '%s'" % name)
                h["time"] = math.inf
                return changed
            if h["ncomp"] != ncomp:
                self.warnings.append("components")
                h["ncomp"] = ncomp
            if h["nent"] != len(ids):
                self.warnings.append("entities")
            h["actual"] = True
        return changed

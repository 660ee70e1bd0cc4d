// gle_io.cpp -- alignment ingestion and site-likelihood output.
//
// Reads the character data GARLI reads through NCL (GarliReader,
// garlireader.cpp; SequenceData::CreateMatrixFromNCL, sequencedata.cpp):
// NEXUS DATA / CHARACTERS blocks (interleaved or not, MISSING / GAP /
// MATCHCHAR symbols, {..} and (..) state sets, quoted and underscore names,
// a TAXA block's labels), the default (starred) EXSET and WTSET of an
// ASSUMPTIONS block, relaxed PHYLIP (a first line per taxon with its name,
// further lines continuing the taxa in turn: interleaved blocks) and FASTA.
// Rows keep the file's symbols (upper-cased, matchchar resolved, a state set
// turned into its IUPAC code for nucleotides); each character carries a
// weight: the WTSET weight, or 0 when the EXSET excludes it.
//
// Writes <ofprefix>.sitelikes.log as Tree::OutputSiteLikelihoods and the
// score-only loop of Population write it (tree.cpp:5501-5565,
// population.cpp:3459-3478, 3702-3727): a header, one line per character
// (site number and -lnL, or "-" for a character that is excluded or carries
// no data), then the tree's line with its -lnL.
//
// Reads the trees of a NEXUS TREES block (a starting tree file) with its
// TRANSLATE table and GarliScore / GarliModel comments.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "gle_internal.h"

namespace gle {

struct Alignment {
  std::vector<std::string> names, rows;
  std::vector<int> weights;
  int datatype = GLE_DATA_DNA;
};

static std::mutex g_aln_mu;
static std::vector<std::unique_ptr<Alignment>> g_alns;

static std::string upper(std::string s) {
  for (char &c : s) c = char(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

// IUPAC code of a set of nucleotide symbols ("AG" -> 'R'); '?' if not DNA.
static char iupac_of_set(const std::string &set) {
  int mask = 0;
  for (char c : set) {
    const int d = nuc_datum(c);
    if (d < 0) return '?';
    mask |= d;
  }
  static const char kCode[16] = {'-', 'A', 'C', 'M', 'G', 'R', 'S', 'V', 'T', 'W', 'Y', 'H', 'K', 'D', 'B', 'N'};
  return kCode[mask & 15];
}

// ------------------------------------------------------------------ NEXUS
class NexusText {
 public:
  explicit NexusText(std::string t) : s_(std::move(t)) {}
  bool eof() {
    skip();
    return p_ >= s_.size();
  }
  // next token: a quoted word, a run of word characters, or one punctuation
  // character; comments are skipped
  std::string token() {
    skip();
    if (p_ >= s_.size()) return "";
    const char c = s_[p_];
    if (c == '\'') return quoted();
    if (std::strchr(";=,(){}:*\"", c)) {
      ++p_;
      return std::string(1, c);
    }
    size_t b = p_;
    while (p_ < s_.size() && !std::isspace(static_cast<unsigned char>(s_[p_])) && !std::strchr(";=,(){}:*\"[", s_[p_]))
      ++p_;
    return s_.substr(b, p_ - b);
  }
  std::string peek() {
    const size_t save = p_;
    std::string t = token();
    p_ = save;
    return t;
  }
  // a taxon name: quoted as written, else underscores become blanks
  std::string name() {
    skip();
    if (p_ < s_.size() && s_[p_] == '\'') return quoted();
    std::string t = token();
    std::replace(t.begin(), t.end(), '_', ' ');
    return t;
  }
  // the rest of the statement up to ';' (consumed), as tokens
  std::vector<std::string> statement() {
    std::vector<std::string> out;
    for (std::string t = token(); !t.empty() && t != ";"; t = token()) out.push_back(t);
    return out;
  }
  // sequence symbols up to the end of the line (interleaved) or until n
  // symbols were read (n > 0); comments and blanks skipped
  std::string symbols(size_t n, bool to_eol) {
    std::string out;
    while (p_ < s_.size()) {
      const char c = s_[p_];
      if (c == '\n' || c == '\r') {
        if (to_eol) {
          if (!out.empty()) break;
          ++p_;
          continue;
        }
        ++p_;
        continue;
      }
      if (std::isspace(static_cast<unsigned char>(c))) {
        ++p_;
        continue;
      }
      if (c == '[') {
        comment();
        continue;
      }
      if (c == ';') break;
      if (n && out.size() >= n) break;
      if (c == '{' || c == '(') {
        const char close = c == '{' ? '}' : ')';
        std::string set;
        ++p_;
        while (p_ < s_.size() && s_[p_] != close) {
          if (!std::isspace(static_cast<unsigned char>(s_[p_])) && s_[p_] != ',') set += s_[p_];
          ++p_;
        }
        ++p_;
        out += iupac_of_set(upper(set));
        continue;
      }
      out += c;
      ++p_;
    }
    return out;
  }
 private:
  void comment() {
    int depth = 0;
    do {
      if (s_[p_] == '[') depth++;
      else if (s_[p_] == ']') depth--;
      ++p_;
    } while (p_ < s_.size() && depth > 0);
  }
  void skip() {
    while (p_ < s_.size()) {
      if (std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
      else if (s_[p_] == '[') comment();
      else break;
    }
  }
  std::string quoted() {
    std::string out;
    ++p_;
    while (p_ < s_.size()) {
      if (s_[p_] == '\'') {
        if (p_ + 1 < s_.size() && s_[p_ + 1] == '\'') {
          out += '\'';
          p_ += 2;
          continue;
        }
        ++p_;
        break;
      }
      out += s_[p_++];
    }
    return out;
  }
  std::string s_;
  size_t p_ = 0;
};

// "key = value" pairs of a DIMENSIONS / FORMAT statement (upper-cased keys)
static std::map<std::string, std::string> settings(const std::vector<std::string> &toks) {
  std::map<std::string, std::string> out;
  for (size_t i = 0; i < toks.size(); i++) {
    const std::string k = upper(toks[i]);
    if (i + 2 < toks.size() && toks[i + 1] == "=") {
      out[k] = toks[i + 2];
      i += 2;
    } else {
      out[k] = "";
    }
  }
  return out;
}

// Character list "1-749 751-1008 1930-. 4" (1-based, '.' = last) -> indices.
static int char_list(const std::vector<std::string> &toks, size_t b, size_t e, int nchar, std::vector<int> &out) {
  auto num = [&](const std::string &t, int &v) {
    if (t == ".") {
      v = nchar;
      return true;
    }
    char *end = nullptr;
    const long x = std::strtol(t.c_str(), &end, 10);
    if (!end || *end || x < 1 || x > nchar) return false;
    v = int(x);
    return true;
  };
  for (size_t i = b; i < e; i++) {
    const std::string &t = toks[i];
    const size_t dash = t.find('-', 1);
    int a, z;
    if (dash != std::string::npos) {
      if (!num(t.substr(0, dash), a) || !num(t.substr(dash + 1), z)) return -1;
    } else if (i + 2 < e && toks[i + 1] == "-") {  // "1 - 5" written apart
      if (!num(t, a) || !num(toks[i + 2], z)) return -1;
      i += 2;
    } else {
      if (!num(t, a)) return -1;
      z = a;
    }
    for (int k = a; k <= z; k++) out.push_back(k - 1);
  }
  return 0;
}

static int read_nexus(const std::string &text, Alignment &A) {
  NexusText tx(text);
  if (upper(tx.token()) != "#NEXUS") return set_error(GLE_ERR_ARG, "not a NEXUS file");
  std::vector<std::string> taxa;
  std::vector<std::string> excluded_sets, weight_sets;  // starred statements, applied after the matrix
  while (!tx.eof()) {
    if (upper(tx.token()) != "BEGIN") continue;
    const std::string block = upper(tx.token());
    tx.statement();
    if (block == "TAXA") {
      for (;;) {
        const std::string cmd = upper(tx.token());
        if (cmd.empty() || cmd == "END" || cmd == "ENDBLOCK") break;
        if (cmd == "TAXLABELS") {
          while (tx.peek() != ";" && !tx.peek().empty()) taxa.push_back(tx.name());
          tx.token();
        } else {
          tx.statement();
        }
      }
      if (upper(tx.peek()) == ";") tx.token();
    } else if (block == "DATA" || block == "CHARACTERS") {
      int ntax = int(taxa.size()), nchar = 0;
      char missing = '?', gap = '-', match = 0;
      bool interleave = false;
      for (;;) {
        const std::string cmd = upper(tx.token());
        if (cmd.empty()) return set_error(GLE_ERR_ARG, "NEXUS: unterminated %s block", block.c_str());
        if (cmd == "END" || cmd == "ENDBLOCK") break;
        if (cmd == "DIMENSIONS") {
          auto st = settings(tx.statement());
          if (st.count("NTAX")) ntax = std::atoi(st["NTAX"].c_str());
          if (st.count("NCHAR")) nchar = std::atoi(st["NCHAR"].c_str());
        } else if (cmd == "FORMAT") {
          auto st = settings(tx.statement());
          if (st.count("DATATYPE")) {
            const std::string d = upper(st["DATATYPE"]);
            A.datatype = (d == "PROTEIN") ? GLE_DATA_PROTEIN
                         : (d == "DNA" || d == "RNA" || d == "NUCLEOTIDE") ? GLE_DATA_DNA
                                                                           : GLE_DATA_OTHER;
          }
          if (st.count("MISSING") && !st["MISSING"].empty()) missing = st["MISSING"][0];
          if (st.count("GAP") && !st["GAP"].empty()) gap = st["GAP"][0];
          if (st.count("MATCHCHAR") && !st["MATCHCHAR"].empty()) match = st["MATCHCHAR"][0];
          if (st.count("INTERLEAVE")) interleave = upper(st["INTERLEAVE"]) != "NO";
        } else if (cmd == "MATRIX") {
          if (ntax <= 0 || nchar <= 0) return set_error(GLE_ERR_ARG, "NEXUS: MATRIX before DIMENSIONS");
          A.names.assign(ntax, "");
          A.rows.assign(ntax, "");
          std::map<std::string, int> index;
          for (int i = 0; i < int(taxa.size()) && i < ntax; i++) index[taxa[i]] = i;
          int next = 0;
          while (upper(tx.peek()) != ";") {
            if (tx.eof()) return set_error(GLE_ERR_ARG, "NEXUS: unterminated MATRIX");
            const std::string nm = tx.name();
            int i;
            auto it = index.find(nm);
            if (it != index.end()) {
              i = it->second;
            } else {
              if (next >= ntax) return set_error(GLE_ERR_ARG, "NEXUS: more taxa than NTAX (%s)", nm.c_str());
              i = next++;
              index[nm] = i;
            }
            A.names[i] = nm;
            std::string &row = A.rows[i];
            row += interleave ? tx.symbols(0, true) : tx.symbols(size_t(nchar) - row.size(), false);
          }
          tx.token();  // ';'
          for (int i = 0; i < ntax; i++) {
            std::string &row = A.rows[i];
            if (int(row.size()) != nchar)
              return set_error(GLE_ERR_ARG, "NEXUS: taxon %s has %d characters, NCHAR is %d", A.names[i].c_str(),
                               int(row.size()), nchar);
            for (int k = 0; k < nchar; k++) {
              char c = row[k];
              if (match && c == match) {
                if (i == 0) return set_error(GLE_ERR_ARG, "NEXUS: matchchar in the first taxon");
                c = A.rows[0][k];
              }
              if (c == missing) c = '?';
              else if (c == gap) c = '-';
              row[k] = char(std::toupper(static_cast<unsigned char>(c)));
            }
          }
        } else {
          tx.statement();
        }
      }
      if (upper(tx.peek()) == ";") tx.token();
    } else if (block == "ASSUMPTIONS" || block == "SETS") {
      for (;;) {
        const std::string cmd = upper(tx.token());
        if (cmd.empty() || cmd == "END" || cmd == "ENDBLOCK") break;
        std::vector<std::string> st = tx.statement();
        if ((cmd == "EXSET" || cmd == "WTSET") && !st.empty() && st[0] == "*") {
          std::string joined;
          for (auto &t : st) joined += t + "\x1f";
          (cmd == "EXSET" ? excluded_sets : weight_sets).push_back(joined);
        }
      }
      if (upper(tx.peek()) == ";") tx.token();
    } else {
      for (std::string t = upper(tx.token()); !t.empty() && t != "END" && t != "ENDBLOCK"; t = upper(tx.token())) {
      }
      if (upper(tx.peek()) == ";") tx.token();
    }
  }
  if (A.rows.empty()) return set_error(GLE_ERR_ARG, "NEXUS: no character matrix");
  const int nchar = int(A.rows[0].size());
  A.weights.assign(nchar, 1);
  auto split = [](const std::string &j) {
    std::vector<std::string> v;
    size_t b = 0;
    for (size_t e; (e = j.find('\x1f', b)) != std::string::npos; b = e + 1) v.push_back(j.substr(b, e - b));
    return v;
  };
  for (const std::string &j : weight_sets) {  // wtset * name [(vector)] = w: list, w: list ...
    std::vector<std::string> t = split(j);
    size_t eq = std::find(t.begin(), t.end(), "=") - t.begin();
    if (eq >= t.size()) return set_error(GLE_ERR_ARG, "NEXUS: bad WTSET");
    bool vector = false;
    for (size_t k = 0; k < eq; k++) vector |= upper(t[k]) == "VECTOR";
    if (vector) {
      if (t.size() - eq - 1 != size_t(nchar)) return set_error(GLE_ERR_ARG, "NEXUS: WTSET vector length");
      for (int k = 0; k < nchar; k++) A.weights[k] = std::atoi(t[eq + 1 + k].c_str());
      continue;
    }
    size_t i = eq + 1;
    while (i < t.size()) {
      if (i + 1 >= t.size() || t[i + 1] != ":") return set_error(GLE_ERR_ARG, "NEXUS: bad WTSET near %s", t[i].c_str());
      const int w = std::atoi(t[i].c_str());
      size_t e = i + 2;
      while (e < t.size() && t[e] != ",") e++;
      std::vector<int> sites;
      if (char_list(t, i + 2, e, nchar, sites)) return set_error(GLE_ERR_ARG, "NEXUS: bad WTSET character list");
      for (int s : sites) A.weights[s] = w;
      i = e + 1;
    }
  }
  for (const std::string &j : excluded_sets) {  // exset * name = list
    std::vector<std::string> t = split(j);
    size_t eq = std::find(t.begin(), t.end(), "=") - t.begin();
    std::vector<int> sites;
    if (eq >= t.size() || char_list(t, eq + 1, t.size(), nchar, sites))
      return set_error(GLE_ERR_ARG, "NEXUS: bad EXSET");
    for (int s : sites) A.weights[s] = 0;
  }
  return GLE_SUCCESS;
}

// ------------------------------------------------------------- PHYLIP, FASTA
static int guess_type(Alignment &A) {
  bool dna = true;
  for (auto &r : A.rows)
    for (char c : r)
      if (c != '?' && c != '-' && nuc_datum(c) < 0) dna = false;
  A.datatype = dna ? GLE_DATA_DNA : GLE_DATA_PROTEIN;
  return GLE_SUCCESS;
}

static std::string clean(const std::string &s) {
  std::string out;
  for (char c : s)
    if (!std::isspace(static_cast<unsigned char>(c))) out += char(std::toupper(static_cast<unsigned char>(c)));
  return out;
}

static int read_phylip(const std::string &text, Alignment &A) {
  std::istringstream in(text);
  int ntax = 0, nchar = 0;
  if (!(in >> ntax >> nchar) || ntax < 1 || nchar < 1) return set_error(GLE_ERR_ARG, "PHYLIP: bad header");
  std::string line;
  std::getline(in, line);
  std::vector<std::string> lines;
  while (std::getline(in, line)) {
    if (clean(line).empty()) continue;
    lines.push_back(line);
  }
  if (int(lines.size()) < ntax) return set_error(GLE_ERR_ARG, "PHYLIP: %d taxa expected", ntax);
  A.names.resize(ntax);
  A.rows.resize(ntax);
  size_t li = 0;
  for (int i = 0; i < ntax; i++, li++) {  // relaxed: the name is the first word
    std::istringstream ls(lines[li]);
    ls >> A.names[i];
    std::string rest;
    std::getline(ls, rest);
    A.rows[i] = clean(rest);
  }
  for (int i = 0; li < lines.size(); li++, i = (i + 1) % ntax) A.rows[i] += clean(lines[li]);  // interleaved
  for (int i = 0; i < ntax; i++)
    if (int(A.rows[i].size()) != nchar)
      return set_error(GLE_ERR_ARG, "PHYLIP: taxon %s has %d characters, not %d", A.names[i].c_str(),
                       int(A.rows[i].size()), nchar);
  A.weights.assign(nchar, 1);
  return guess_type(A);
}

static int read_fasta(const std::string &text, Alignment &A) {
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '>') {
      std::string nm = line.substr(1);
      nm.erase(0, nm.find_first_not_of(" \t"));
      nm.erase(nm.find_last_not_of(" \t\r") + 1);
      A.names.push_back(nm);
      A.rows.emplace_back();
    } else if (!A.rows.empty()) {
      A.rows.back() += clean(line);
    }
  }
  if (A.rows.empty()) return set_error(GLE_ERR_ARG, "FASTA: no sequences");
  for (auto &r : A.rows)
    if (r.size() != A.rows[0].size()) return set_error(GLE_ERR_ARG, "FASTA: sequences differ in length");
  A.weights.assign(A.rows[0].size(), 1);
  return guess_type(A);
}

static Alignment *get_aln(int h) {
  std::lock_guard<std::mutex> lk(g_aln_mu);
  if (h < 0 || h >= int(g_alns.size())) return nullptr;
  return g_alns[h].get();
}

}  // namespace gle

using namespace gle;

extern "C" int gle_alignment_read(const char *path, int *handle) {
  if (!path || !handle) return set_error(GLE_ERR_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return set_error(GLE_ERR_ARG, "cannot open %s", path);
  std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  auto A = std::make_unique<Alignment>();
  size_t b = text.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return set_error(GLE_ERR_ARG, "%s is empty", path);
  int rc;
  if (text[b] == '>') rc = read_fasta(text, *A);
  else if (upper(text.substr(b, 6)) == "#NEXUS") rc = read_nexus(text.substr(b), *A);
  else rc = read_phylip(text, *A);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_aln_mu);
  g_alns.push_back(std::move(A));
  *handle = int(g_alns.size()) - 1;
  return GLE_SUCCESS;
}

extern "C" int gle_alignment_dims(int h, int *ntax, int *nchar, int *datatype) {
  Alignment *A = get_aln(h);
  if (!A || !ntax || !nchar || !datatype) return set_error(GLE_ERR_ARG, "bad alignment handle or argument");
  *ntax = int(A->rows.size());
  *nchar = A->rows.empty() ? 0 : int(A->rows[0].size());
  *datatype = A->datatype;
  return GLE_SUCCESS;
}

extern "C" int gle_alignment_taxon(int h, int i, char *name, int cap) {
  Alignment *A = get_aln(h);
  if (!A || !name || cap < 1 || i < 0 || i >= int(A->names.size())) return set_error(GLE_ERR_ARG, "bad argument");
  std::snprintf(name, size_t(cap), "%s", A->names[i].c_str());
  return int(A->names[i].size());
}

extern "C" int gle_alignment_row(int h, int i, char *row) {
  Alignment *A = get_aln(h);
  if (!A || !row || i < 0 || i >= int(A->rows.size())) return set_error(GLE_ERR_ARG, "bad argument");
  std::memcpy(row, A->rows[i].data(), A->rows[i].size());
  return GLE_SUCCESS;
}

extern "C" int gle_alignment_weights(int h, int *weights) {
  Alignment *A = get_aln(h);
  if (!A || !weights) return set_error(GLE_ERR_ARG, "bad argument");
  std::copy(A->weights.begin(), A->weights.end(), weights);
  return GLE_SUCCESS;
}

extern "C" int gle_alignment_free(int h) {
  std::lock_guard<std::mutex> lk(g_aln_mu);
  if (h < 0 || h >= int(g_alns.size()) || !g_alns[h]) return set_error(GLE_ERR_ARG, "bad alignment handle");
  g_alns[h].reset();
  return GLE_SUCCESS;
}

// ------------------------------------------------------ state matrices

extern "C" int gle_alignment_states(int h, unsigned char *out) {
  Alignment *A = get_aln(h);
  if (!A || !out) return set_error(GLE_ERR_ARG, "bad argument");
  const bool dna = A->datatype == GLE_DATA_DNA;
  if (!dna && A->datatype != GLE_DATA_PROTEIN) return set_error(GLE_ERR_UNSUPPORTED, "neither DNA nor protein data");
  const size_t nc = A->rows.empty() ? 0 : A->rows[0].size();
  for (size_t t = 0; t < A->rows.size(); t++)
    for (size_t j = 0; j < nc; j++) {
      const char c = A->rows[t][j];
      const int d = dna ? nuc_datum(c) : aa_datum(c);
      if (d < 0) return set_error(GLE_ERR_ARG, "symbol '%c' of taxon %zu, character %zu", c, t + 1, j + 1);
      out[t * nc + j] = (unsigned char)d;
    }
  return GLE_SUCCESS;
}

// The three nucleotide bit sets of codon k of taxon t.
static int codon_masks(const Alignment &A, size_t t, size_t k, int m[3]) {
  for (int i = 0; i < 3; i++) {
    const char c = A.rows[t][3 * k + i];
    m[i] = nuc_datum(c);
    if (m[i] < 0) return set_error(GLE_ERR_ARG, "symbol '%c' of taxon %zu, character %zu", c, t + 1, 3 * k + i + 1);
  }
  return GLE_SUCCESS;
}

static int codon_dims(Alignment *A, size_t *ncod) {
  if (!A) return set_error(GLE_ERR_ARG, "bad alignment handle");
  if (A->datatype != GLE_DATA_DNA) return set_error(GLE_ERR_UNSUPPORTED, "codons need DNA data");
  const size_t nc = A->rows.empty() ? 0 : A->rows[0].size();
  if (nc % 3) return set_error(GLE_ERR_ARG, "%zu nucleotides is not a whole number of codons", nc);
  *ncod = nc / 3;
  return GLE_SUCCESS;
}

static const char kBase[4] = {'A', 'C', 'G', 'T'};

// CodonData::FillCodonMatrixFromDNA (sequencedata.cpp:364-470), standard
// code.  Only single-base sets name a base (the reference's bit-set to index
// table); a codon with any other set is missing.  Base counts at each codon
// position are taken over the unambiguous codons, stops included, before a
// stop is refused or turned into missing data.
extern "C" int gle_alignment_codons(int h, int ignore_stops, unsigned char *out, double *base_freqs) {
  Alignment *A = get_aln(h);
  size_t ncod;
  int rc = codon_dims(A, &ncod);
  if (rc) return rc;
  if (!out) return set_error(GLE_ERR_ARG, "null output");
  int sense[64], n = 0;
  for (int c = 0; c < 64; c++) sense[c] = standard_code_aa(c) == 20 ? -1 : n++;
  double pos[3][4] = {{0}}, total = 0.0;
  for (size_t t = 0; t < A->rows.size(); t++)
    for (size_t k = 0; k < ncod; k++) {
      int m[3], b[3];
      if ((rc = codon_masks(*A, t, k, m))) return rc;
      bool whole = true;
      for (int i = 0; i < 3; i++) {
        whole = whole && __builtin_popcount(unsigned(m[i])) == 1;
        b[i] = __builtin_ctz(unsigned(m[i]) | 16u);
      }
      unsigned char v = 61;
      if (whole) {
        for (int i = 0; i < 3; i++) pos[i][b[i]] += 1.0;
        total += 1.0;
        const int c = 16 * b[0] + 4 * b[1] + b[2];
        if (sense[c] >= 0) v = (unsigned char)sense[c];
        else if (!ignore_stops)
          return set_error(GLE_ERR_ARG, "stop codon %c%c%c at codon %zu (nucleotide %zu) of taxon %zu", kBase[b[0]],
                           kBase[b[1]], kBase[b[2]], k + 1, 3 * k + 1, t + 1);
      }
      out[t * ncod + k] = v;
    }
  if (base_freqs)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 4; j++) base_freqs[4 * i + j] = total > 0.0 ? pos[i][j] / total : 0.0;
  return GLE_SUCCESS;
}

// AminoacidData::FillAminoacidMatrixFromDNA (sequencedata.cpp:473-590),
// standard code.  A codon whose first two positions are both fully missing
// is missing (the reference tests the sum of positions 1, 2 and 1 against
// three missing codes).  Otherwise every resolution of its bit sets is
// translated in A, C, G, T order: resolutions to two different amino acids
// make it missing; a stop among resolutions that also give an amino acid
// makes it missing; resolutions that are all stops are an error, and with
// ignore_stops any stop makes the codon missing.
extern "C" int gle_alignment_translate(int h, int ignore_stops, unsigned char *out) {
  Alignment *A = get_aln(h);
  size_t ncod;
  int rc = codon_dims(A, &ncod);
  if (rc) return rc;
  if (!out) return set_error(GLE_ERR_ARG, "null output");
  for (size_t t = 0; t < A->rows.size(); t++)
    for (size_t k = 0; k < ncod; k++) {
      int m[3];
      if ((rc = codon_masks(*A, t, k, m))) return rc;
      int prot = -1, first_stop = -1;
      if (2 * m[0] + m[1] == 45) {
        prot = 20;
      } else {
        bool done = false;
        for (int x = 0; x < 4 && !done; x++)
          for (int y = 0; y < 4 && !done; y++)
            for (int z = 0; z < 4 && !done; z++) {
              if (!(m[0] >> x & 1) || !(m[1] >> y & 1) || !(m[2] >> z & 1)) continue;
              const int c = 16 * x + 4 * y + z, a = standard_code_aa(c);
              if (a == 20) {
                if (first_stop < 0) first_stop = c;
                if (ignore_stops) {
                  prot = 20;
                  done = true;
                }
              } else if (prot < 0) {
                prot = a;
              } else if (a != prot) {
                prot = 20;
                done = true;
              }
            }
        if (first_stop >= 0 && !done) {
          if (prot < 0)
            return set_error(GLE_ERR_ARG, "stop codon %c%c%c at codon %zu (nucleotide %zu) of taxon %zu",
                             kBase[first_stop >> 4], kBase[(first_stop >> 2) & 3], kBase[first_stop & 3], k + 1,
                             3 * k + 1, t + 1);
          prot = 20;
        }
      }
      out[t * ncod + k] = (unsigned char)prot;
    }
  return GLE_SUCCESS;
}

extern "C" int gle_write_sitelikes(const char *path, int append, int tree_num, double lnl, int nchar,
                                   const double *site_lnl, const int *included) {
  if (!path || nchar < 0 || (nchar > 0 && (!site_lnl || !included))) return set_error(GLE_ERR_ARG, "bad argument");
  std::FILE *f = std::fopen(path, append ? "a" : "w");
  if (!f) return set_error(GLE_ERR_ARG, "cannot write %s", path);
  if (!append) std::fputs("Tree\t-lnL\tSite\t-lnL\n", f);
  for (int s = 0; s < nchar; s++) {
    if (included[s]) std::fprintf(f, "\t\t%d\t%.8f\n", s + 1, -site_lnl[s]);
    else std::fprintf(f, "\t\t%d\t-\n", s + 1);
  }
  std::fprintf(f, "%d\t%.12g\n", tree_num, -lnl);
  const bool ok = std::fclose(f) == 0;
  return ok ? GLE_SUCCESS : set_error(GLE_ERR_ARG, "write to %s failed", path);
}

// Tree::OutputSiteDerivatives (tree.cpp:5567-5595): an ostream at precision
// 10 in its default float format, i.e. %.10g.
extern "C" int gle_write_site_derivatives(const char *ordered_path, const char *packed_path, int subset, int nchar,
                                          const int *site_pattern, int npat, const double *lnl, const double *d1,
                                          const double *d2, const int *u1, const int *u2) {
  if (nchar < 0 || npat < 0 || (npat > 0 && (!lnl || !d1 || !d2 || !u1)) || (nchar > 0 && !site_pattern))
    return set_error(GLE_ERR_ARG, "bad argument");
  for (int c = 0; c < nchar; c++)
    if (site_pattern[c] < -1 || site_pattern[c] >= npat) return set_error(GLE_ERR_ARG, "character %d: bad pattern", c);
  auto row = [&](std::FILE *f, int label, int q) {
    std::fprintf(f, "%d\t%.10g\t%.10g\t%.10g\t%d", label, lnl[q], d1[q], d2[q], u1[q]);
    if (u2) std::fprintf(f, "\t%d\n", u2[q]);
    else std::fputs("\t-\n", f);
  };
  if (ordered_path) {
    std::FILE *f = std::fopen(ordered_path, "w");
    if (!f) return set_error(GLE_ERR_ARG, "cannot write %s", ordered_path);
    std::fprintf(f, "Partition subset %d\nsite#\ttruelnL\td1\td2\tunder1\tunder2\n", subset);
    for (int c = 0; c < nchar; c++) {
      if (site_pattern[c] < 0) std::fprintf(f, "%d\tgap\t-\t-\t-\t-", c + 1);
      else row(f, c + 1, site_pattern[c]);
    }
    if (std::fclose(f)) return set_error(GLE_ERR_ARG, "write to %s failed", ordered_path);
  }
  if (packed_path) {
    std::FILE *f = std::fopen(packed_path, "w");
    if (!f) return set_error(GLE_ERR_ARG, "cannot write %s", packed_path);
    std::fprintf(f, "Partition subset %d\npackedIndex\ttruelnL\td1\td2\tunder1\tunder2\n", subset);
    for (int q = 0; q < npat; q++) row(f, q, q);
    if (std::fclose(f)) return set_error(GLE_ERR_ARG, "write to %s failed", packed_path);
  }
  return GLE_SUCCESS;
}


// --------------------------------------------------------------- start trees
// The TREES blocks of a NEXUS file as GarliReader reads them through NCL for
// a starting tree (garlireader.cpp:793, 1245; Population::
// SeedPopulationWithStartingTree, population.cpp:1186): a TRANSLATE table
// and tree statements whose comments may carry the tree's [&U] flag, its
// [!GarliScore s] and its [!GarliModel ...] parameters.
namespace gle {

struct TreeFile {
  std::map<std::string, std::string> translate;  // token -> taxon label
  struct Entry {
    std::string name, text;  // the newick text after '=' (comments kept)
  };
  std::vector<Entry> trees;
};

static std::mutex g_tf_mu;
static std::vector<std::unique_ptr<TreeFile>> g_tfs;

static bool ieq_at(const std::string &s, size_t p, const char *w) {
  const size_t n = std::strlen(w);
  if (p + n > s.size()) return false;
  for (size_t i = 0; i < n; i++)
    if (std::tolower(static_cast<unsigned char>(s[p + i])) != w[i]) return false;
  return true;
}

// A newick string: nested children, labels (quoted or bare, '_' read as a
// blank), ":length", [comments] anywhere.
struct NewickNode {
  std::string label;
  double len = -1.0;  // -1: not given
  std::vector<int> kids;
};

class NewickParser {
 public:
  explicit NewickParser(const std::string &s) : s_(s) {}
  int parse(std::vector<NewickNode> &out, std::string &err) {
    out.clear();
    const int root = node(out, err);
    if (!err.empty()) return -1;
    skip();
    if (p_ < s_.size() && s_[p_] == ';') ++p_;
    skip();
    if (p_ != s_.size()) {
      err = "text after the tree";
      return -1;
    }
    return root;
  }

 private:
  void skip() {
    while (p_ < s_.size()) {
      if (std::isspace(static_cast<unsigned char>(s_[p_]))) {
        ++p_;
      } else if (s_[p_] == '[') {
        int depth = 0;
        do {
          if (s_[p_] == '[') depth++;
          else if (s_[p_] == ']') depth--;
          ++p_;
        } while (p_ < s_.size() && depth > 0);
      } else {
        break;
      }
    }
  }
  std::string label() {
    skip();
    std::string out;
    if (p_ < s_.size() && s_[p_] == '\'') {
      ++p_;
      while (p_ < s_.size()) {
        if (s_[p_] == '\'') {
          if (p_ + 1 < s_.size() && s_[p_ + 1] == '\'') {
            out += '\'';
            p_ += 2;
            continue;
          }
          ++p_;
          break;
        }
        out += s_[p_++];
      }
      return out;
    }
    while (p_ < s_.size() && !std::strchr("(),:;[", s_[p_]) && !std::isspace(static_cast<unsigned char>(s_[p_])))
      out += s_[p_++];
    std::replace(out.begin(), out.end(), '_', ' ');
    return out;
  }
  int node(std::vector<NewickNode> &out, std::string &err) {
    skip();
    NewickNode nd;
    if (p_ < s_.size() && s_[p_] == '(') {
      ++p_;
      for (;;) {
        const int c = node(out, err);
        if (!err.empty()) return -1;
        nd.kids.push_back(c);
        skip();
        if (p_ < s_.size() && s_[p_] == ',') {
          ++p_;
          continue;
        }
        if (p_ < s_.size() && s_[p_] == ')') {
          ++p_;
          break;
        }
        err = "expected ',' or ')'";
        return -1;
      }
    }
    nd.label = label();
    skip();
    if (p_ < s_.size() && s_[p_] == ':') {
      ++p_;
      skip();
      const char *b = s_.c_str() + p_;
      char *e = nullptr;
      nd.len = std::strtod(b, &e);
      if (e == b) {
        err = "bad branch length";
        return -1;
      }
      p_ += size_t(e - b);
    }
    if (nd.kids.empty() && nd.label.empty()) {
      err = "a tip without a label";
      return -1;
    }
    out.push_back(nd);
    return int(out.size()) - 1;
  }
  const std::string &s_;
  size_t p_ = 0;
};

static int read_tree_file(const std::string &text, TreeFile &F) {
  size_t p = 0;
  const size_t n = text.size();
  auto skip_ws = [&] {
    while (p < n && std::isspace(static_cast<unsigned char>(text[p]))) ++p;
  };
  auto skip_comment = [&] {
    int depth = 0;
    do {
      if (text[p] == '[') depth++;
      else if (text[p] == ']') depth--;
      ++p;
    } while (p < n && depth > 0);
  };
  // the end of a statement: the next ';' outside comments and quotes
  auto statement_end = [&](size_t from) {
    size_t q = from;
    while (q < n) {
      if (text[q] == '[') {
        int depth = 0;
        do {
          if (text[q] == '[') depth++;
          else if (text[q] == ']') depth--;
          ++q;
        } while (q < n && depth > 0);
        continue;
      }
      if (text[q] == '\'') {
        ++q;
        while (q < n && text[q] != '\'') ++q;
        ++q;
        continue;
      }
      if (text[q] == ';') return q;
      ++q;
    }
    return n;
  };
  bool in_trees = false;
  while (p < n) {
    skip_ws();
    if (p >= n) break;
    if (text[p] == '[') {
      skip_comment();
      continue;
    }
    if (!in_trees) {
      if (ieq_at(text, p, "#nexus")) {
        p += 6;
        continue;
      }
      if (ieq_at(text, p, "begin")) {
        p += 5;
        skip_ws();
        const bool trees = ieq_at(text, p, "trees");
        p = statement_end(p) + 1;
        in_trees = trees;
        continue;
      }
      p = statement_end(p) + 1;
      continue;
    }
    if (ieq_at(text, p, "endblock") || ieq_at(text, p, "end")) {
      in_trees = false;
      p = statement_end(p) + 1;
      continue;
    }
    if (ieq_at(text, p, "translate")) {
      const size_t e = statement_end(p);
      std::string body = text.substr(p + 9, e - p - 9);
      size_t b = 0;
      while (b <= body.size()) {
        size_t c = body.find(',', b);
        if (c == std::string::npos) c = body.size();
        std::istringstream is(body.substr(b, c - b));
        std::string tok, lab;
        if (is >> tok) {
          std::getline(is, lab);
          const size_t s0 = lab.find_first_not_of(" \t\r\n"), s1 = lab.find_last_not_of(" \t\r\n");
          lab = s0 == std::string::npos ? "" : lab.substr(s0, s1 - s0 + 1);
          if (lab.size() >= 2 && lab.front() == '\'' && lab.back() == '\'') lab = lab.substr(1, lab.size() - 2);
          else std::replace(lab.begin(), lab.end(), '_', ' ');
          F.translate[tok] = lab;
        }
        b = c + 1;
      }
      p = e + 1;
      continue;
    }
    if (ieq_at(text, p, "utree") || ieq_at(text, p, "tree")) {
      const size_t e = statement_end(p);
      const std::string st = text.substr(p, e - p);
      const size_t eq = st.find('=');
      if (eq == std::string::npos) return set_error(GLE_ERR_ARG, "tree statement without '='");
      std::string name = st.substr(ieq_at(st, 0, "utree") ? 5 : 4, eq - (ieq_at(st, 0, "utree") ? 5 : 4));
      const size_t s0 = name.find_first_not_of(" \t\r\n*"), s1 = name.find_last_not_of(" \t\r\n");
      name = s0 == std::string::npos ? "" : name.substr(s0, s1 - s0 + 1);
      F.trees.push_back({name, st.substr(eq + 1)});
      p = e + 1;
      continue;
    }
    p = statement_end(p) + 1;
  }
  if (F.trees.empty()) return set_error(GLE_ERR_ARG, "no tree statement in a TREES block");
  return GLE_SUCCESS;
}

static TreeFile *get_tf(int h) {
  std::lock_guard<std::mutex> lk(g_tf_mu);
  if (h < 0 || h >= int(g_tfs.size()) || !g_tfs[h]) return nullptr;
  return g_tfs[h].get();
}

// The value after "!key" in the tree's comments ("" if absent).
static std::string comment_value(const std::string &text, const char *key) {
  const std::string k = std::string("!") + key;
  size_t b = 0;
  while ((b = text.find('[', b)) != std::string::npos) {
    const size_t e = text.find(']', b);
    if (e == std::string::npos) break;
    std::string c = text.substr(b + 1, e - b - 1);
    size_t s = c.find_first_not_of(" \t\r\n");
    if (s != std::string::npos && c.compare(s, k.size(), k) == 0) {
      c = c.substr(s + k.size());
      const size_t a = c.find_first_not_of(" \t\r\n"), z = c.find_last_not_of(" \t\r\n");
      return a == std::string::npos ? "" : c.substr(a, z - a + 1);
    }
    b = e + 1;
  }
  return "";
}

}  // namespace gle

extern "C" int gle_treefile_read(const char *path, int *handle) {
  if (!path || !handle) return set_error(GLE_ERR_ARG, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) return set_error(GLE_ERR_ARG, "cannot open %s", path);
  const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  auto F = std::make_unique<TreeFile>();
  const int rc = read_tree_file(text, *F);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_tf_mu);
  g_tfs.push_back(std::move(F));
  *handle = int(g_tfs.size()) - 1;
  return GLE_SUCCESS;
}

extern "C" int gle_treefile_count(int h) {
  TreeFile *F = get_tf(h);
  if (!F) return set_error(GLE_ERR_ARG, "bad tree file handle %d", h);
  return int(F->trees.size());
}

extern "C" int gle_treefile_model(int h, int k, char *buf, int cap) {
  TreeFile *F = get_tf(h);
  if (!F || k < 0 || k >= int(F->trees.size()) || !buf || cap < 1) return set_error(GLE_ERR_ARG, "bad argument");
  const std::string v = comment_value(F->trees[k].text, "GarliModel");
  const size_t m = std::min(v.size(), size_t(cap - 1));
  std::memcpy(buf, v.data(), m);
  buf[m] = 0;
  return int(v.size());
}

// Tree k as the tree layer's arrays, following Tree::Tree(const char *, ..)
// (tree.cpp:309-576): internal nodes numbered n_taxa + 1.. in the order the
// description opens them; a missing length is the starting length 0.05
// (DEF_STARTING_BRLEN, defs.h:108) and a given one is held to [1e-8, 100]
// (DEF_MIN_BRLEN / DEF_MAX_BRLEN, defs.h:106-107); a bifurcating top is made
// a trifurcation by MakeTrifurcatingRoot's rule (tree.cpp:761-797, the left
// child dissolved when internal, else the right one, its length added to its
// sibling's) and the dissolved node's number given up (EliminateNode).  A
// polytomy is resolved as ArbitrarilyBifurcate does (tree.cpp:855-923), one
// new node of the starting length per step, but joining the node's first two
// children where the reference draws the pair at random; the return value is
// then 1.
extern "C" int gle_treefile_tree(int h, int k, int n_taxa, const char *const *names, int *parent, double *blen,
                                 double *score) {
  constexpr double kStartLen = 0.05, kMinLen = 1e-8, kMaxLen = 100.0;
  TreeFile *F = get_tf(h);
  if (!F || k < 0 || k >= int(F->trees.size()) || n_taxa < 3 || !parent || !blen)
    return set_error(GLE_ERR_ARG, "bad argument");
  std::vector<NewickNode> nodes;
  std::string err;
  const int root = NewickParser(F->trees[k].text).parse(nodes, err);
  if (root < 0) return set_error(GLE_ERR_ARG, "tree %s: %s", F->trees[k].name.c_str(), err.c_str());
  for (int u = 0; u < int(nodes.size()); u++) {
    if (u == root) continue;
    double &l = nodes[u].len;
    l = l < 0.0 ? kStartLen : std::min(std::max(l, kMinLen), kMaxLen);
    if (nodes[u].kids.size() == 1) return set_error(GLE_ERR_ARG, "an internal node with one child");
  }
  // tip label -> taxon number 1..N: its TRANSLATE label matched by name among
  // the given taxa (a blank matching '_'), else the label itself as a number
  // 1..N (a TRANSLATE table's tokens are the taxa's numbers)
  auto as_number = [&](const std::string &l) -> int {
    char *e = nullptr;
    const long v = std::strtol(l.c_str(), &e, 10);
    return (!l.empty() && e && *e == 0 && v >= 1 && v <= n_taxa) ? int(v) : -1;
  };
  auto taxon = [&](const std::string &lab) -> int {
    std::string l = lab;
    auto it = F->translate.find(lab);
    if (it != F->translate.end()) l = it->second;
    if (names)
      for (int i = 0; i < n_taxa; i++) {
        std::string alt = names[i];
        std::replace(alt.begin(), alt.end(), '_', ' ');
        if (l == names[i] || l == alt) return i + 1;
      }
    const int v = as_number(lab);
    return v > 0 ? v : as_number(l);
  };
  // the node dissolved by a bifurcating top
  int gone = -1;
  if (nodes[root].kids.size() == 2) {
    const int a = nodes[root].kids[0], b = nodes[root].kids[1];
    gone = !nodes[a].kids.empty() ? a : (!nodes[b].kids.empty() ? b : -1);
    if (gone < 0) return set_error(GLE_ERR_ARG, "a two-taxon tree");
  }
  if (nodes[root].kids.size() < 2) return set_error(GLE_ERR_ARG, "the tree's top node has one child");
  // numbers: internal nodes in the order the description opens them (a
  // preorder of the tree as written), tips by taxon
  const int nn = 2 * n_taxa - 2;
  std::vector<int> num(nodes.size(), -1);
  int next = n_taxa + 1;
  {
    std::vector<int> stack = {root};
    while (!stack.empty()) {
      const int u = stack.back();
      stack.pop_back();
      if (u == root) {
        num[u] = 0;
      } else if (nodes[u].kids.empty()) {
        num[u] = taxon(nodes[u].label);
        if (num[u] < 0) return set_error(GLE_ERR_ARG, "tip %s is not a taxon of the data", nodes[u].label.c_str());
      } else if (u != gone) {
        num[u] = next++;
      }
      for (auto c = nodes[u].kids.rbegin(); c != nodes[u].kids.rend(); ++c) stack.push_back(*c);
    }
  }
  if (gone >= 0) {
    const int other = nodes[root].kids[0] == gone ? nodes[root].kids[1] : nodes[root].kids[0];
    nodes[other].len += nodes[gone].len;
    std::vector<int> kids = {other};
    kids.insert(kids.end(), nodes[gone].kids.begin(), nodes[gone].kids.end());
    nodes[root].kids = kids;
  }
  // polytomies: join the first two children under a new node until resolved
  bool resolved = false;
  for (int u = 0; u < int(nodes.size()); u++) {
    if (u == gone || nodes[u].kids.empty()) continue;
    const size_t want = u == root ? 3 : 2;
    while (nodes[u].kids.size() > want) {
      NewickNode j;
      j.len = kStartLen;
      j.kids = {nodes[u].kids[0], nodes[u].kids[1]};
      nodes.push_back(j);
      num.push_back(next++);
      std::vector<int> kids = {int(nodes.size()) - 1};
      kids.insert(kids.end(), nodes[u].kids.begin() + 2, nodes[u].kids.end());
      nodes[u].kids = kids;
      resolved = true;
    }
  }
  if (next != nn) return set_error(GLE_ERR_ARG, "tree has %d internal nodes, %d taxa need %d", next - n_taxa - 1,
                                   n_taxa, n_taxa - 2);
  std::vector<int> par(nn, -2);
  std::vector<double> len(nn, 0.0);
  std::vector<std::pair<int, int>> stack = {{root, -1}};
  while (!stack.empty()) {
    const auto [u, pu] = stack.back();
    stack.pop_back();
    const int me = num[u];
    if (me < 0 || me >= nn) return set_error(GLE_ERR_ARG, "tree does not fit %d taxa", n_taxa);
    if (par[me] != -2) return set_error(GLE_ERR_ARG, "taxon %d appears twice", me);
    par[me] = pu < 0 ? -1 : num[pu];
    len[me] = pu < 0 ? 0.0 : nodes[u].len;
    for (int c : nodes[u].kids) stack.push_back({c, u});
  }
  for (int i = 0; i < nn; i++)
    if (par[i] == -2) return set_error(GLE_ERR_ARG, "the tree does not hold all %d taxa", n_taxa);
  std::copy(par.begin(), par.end(), parent);
  std::copy(len.begin(), len.end(), blen);
  if (score) {
    const std::string v = comment_value(F->trees[k].text, "GarliScore");
    *score = v.empty() ? std::nan("") : std::strtod(v.c_str(), nullptr);
  }
  return resolved ? 1 : GLE_SUCCESS;
}

extern "C" int gle_treefile_free(int h) {
  std::lock_guard<std::mutex> lk(g_tf_mu);
  if (h < 0 || h >= int(g_tfs.size()) || !g_tfs[h]) return set_error(GLE_ERR_ARG, "bad tree file handle %d", h);
  g_tfs[h].reset();
  return GLE_SUCCESS;
}

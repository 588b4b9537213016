// Alignment ingest: PHYLIP / FASTA parsing, sequence-type detection, state coding and
// site-pattern compression (IQ-TREE's Alignment reader, alignment/alignment.cpp).
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "plk.h"
#include "plk_align.h"
#include "plk_internal.h"
#include "plk_model.h"

namespace plk {

struct Align {
    std::vector<std::string> names;
    int ntaxa = 0, nsites = 0, npat = 0, nstates = 0, seq_type = 0, gaps_only = 0;
    std::vector<uint8_t> states;  // ntaxa x npat
    std::vector<double> freq;
    std::vector<int32_t> site_pattern, flags;
};

}  // namespace plk

struct plk_align : public plk::Align {};

namespace plk {
namespace {

struct ParseError {
    std::string msg;
};

std::vector<std::string> split_lines(const char *text, size_t len) {
    std::vector<std::string> lines;
    size_t i = 0;
    while (i < len) {
        size_t j = i;
        while (j < len && text[j] != '\n' && text[j] != '\r') j++;
        lines.emplace_back(text + i, j - i);
        if (j < len && text[j] == '\r' && j + 1 < len && text[j + 1] == '\n') j++;
        i = j + 1;
    }
    return lines;
}

// processSeq (alignment/alignment.cpp:2444-2472)
void process_seq(std::string &seq, const std::string &line, int line_num) {
    for (size_t k = 0; k < line.size(); k++) {
        const char ch = line[k];
        if ((unsigned char)ch <= ' ') continue;
        if (std::isalnum((unsigned char)ch) || ch == '-' || ch == '?' || ch == '.' || ch == '*' || ch == '~') {
            seq.push_back((char)std::toupper((unsigned char)ch));
        } else if (ch == '!') {
            seq.push_back(ch);
        } else if (ch == '(' || ch == '{') {
            while (k < line.size() && line[k] != ')' && line[k] != '}') k++;
            if (k == line.size())
                throw ParseError{"Line " + std::to_string(line_num) + ": No matching close-bracket ) or } found"};
            seq.push_back('?');
        } else {
            throw ParseError{"Line " + std::to_string(line_num) + ": Unrecognized character " + std::string(1, ch)};
        }
    }
}

void read_header(const std::string &line, int &nseq, int &nsite) {
    std::istringstream in(line);
    if (!(in >> nseq >> nsite)) throw ParseError{"Invalid PHYLIP format. First line must contain number of sequences and sites"};
    if (nseq < 3) throw ParseError{"There must be at least 3 sequences"};
    if (nsite < 1) throw ParseError{"No alignment columns"};
}

std::string cut_name(std::string &line) {
    size_t pos = line.find_first_of(" \t");
    if (pos == std::string::npos) pos = 10;  // assume standard phylip
    std::string name = line.substr(0, pos);
    line.erase(0, std::min(pos, line.size()));
    return name;
}

// doReadPhylip (2474-2556): interleaved blocks, sequence ids wrap around
int read_phylip(const std::vector<std::string> &lines, std::vector<std::string> &names, std::vector<std::string> &seqs) {
    int nseq = 0, nsite = 0, seq_id = 0, line_num = 0;
    for (const std::string &raw : lines) {
        line_num++;
        std::string line = raw;
        if (line.empty()) continue;
        if (nseq == 0) {
            read_header(line, nseq, nsite);
            names.assign(nseq, "");
            seqs.assign(nseq, "");
            continue;
        }
        if (names[seq_id].empty()) names[seq_id] = cut_name(line);
        const size_t old_len = seqs[seq_id].size();
        process_seq(seqs[seq_id], line, line_num);
        if (seqs[seq_id].size() != seqs[0].size())
            throw ParseError{"Line " + std::to_string(line_num) + ": Sequence " + names[seq_id] +
                             " has wrong sequence length " + std::to_string(seqs[seq_id].size())};
        if (seqs[seq_id].size() > old_len) seq_id++;
        if (seq_id == nseq) seq_id = 0;
    }
    if (nseq == 0) throw ParseError{"Invalid PHYLIP format. First line must contain number of sequences and sites"};
    return nsite;
}

// doReadPhylipSequential (2567-2633): each sequence runs until it has nsite characters
int read_phylip_sequential(const std::vector<std::string> &lines, std::vector<std::string> &names,
                           std::vector<std::string> &seqs) {
    int nseq = 0, nsite = 0, seq_id = 0, line_num = 0;
    for (const std::string &raw : lines) {
        line_num++;
        std::string line = raw;
        if (line.empty()) continue;
        if (nseq == 0) {
            read_header(line, nseq, nsite);
            names.assign(nseq, "");
            seqs.assign(nseq, "");
            continue;
        }
        if (seq_id >= nseq) throw ParseError{"Line " + std::to_string(line_num) + ": Too many sequences detected"};
        if (names[seq_id].empty()) names[seq_id] = cut_name(line);
        process_seq(seqs[seq_id], line, line_num);
        if ((int)seqs[seq_id].size() > nsite)
            throw ParseError{"Line " + std::to_string(line_num) + ": Sequence " + names[seq_id] + " is too long (" +
                             std::to_string(seqs[seq_id].size()) + ")"};
        if ((int)seqs[seq_id].size() == nsite) seq_id++;
    }
    if (nseq == 0) throw ParseError{"Invalid PHYLIP format. First line must contain number of sequences and sites"};
    return nsite;
}

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t"), b = s.find_last_not_of(" \t");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// doReadFasta (2708-2808), with the name shortening of 2746-2790
int read_fasta(const std::vector<std::string> &lines, std::vector<std::string> &names, std::vector<std::string> &seqs) {
    int line_num = 0;
    for (const std::string &line : lines) {
        line_num++;
        if (line.empty()) continue;
        if (line[0] == '>') {
            names.push_back(trim(line.substr(1)));
            seqs.emplace_back();
            continue;
        }
        if (seqs.empty()) throw ParseError{"First line must begin with '>' to define sequence name"};
        process_seq(seqs.back(), line, line_num);
    }
    std::vector<std::string> shortened(names.size()), remain = names;
    for (int step = 0; step < 4; step++) {
        bool duplicated = false;
        std::unordered_set<std::string> seen;
        for (size_t i = 0; i < names.size(); i++) {
            if (remain[i].empty()) continue;
            size_t pos = remain[i].find_first_of(" \t");
            if (pos == std::string::npos) {
                shortened[i] += remain[i];
                remain[i].clear();
            } else {
                shortened[i] += remain[i].substr(0, pos);
                remain[i] = "_" + remain[i].substr(pos + 1);
            }
            if (!duplicated) duplicated = !seen.insert(shortened[i]).second;
        }
        if (!duplicated) break;
    }
    names = shortened;
    if (seqs.empty()) throw ParseError{"No sequences found"};
    return (int)seqs.front().size();
}

// detectSequenceType (1720-1786): DNA when every letter is a nucleotide code and > 90% are
// proper nucleotides; protein when > 90% of the letters are the 20 amino acids.
int detect_type(const std::vector<std::string> &seqs) {
    size_t proper_nuc = 0, nuc = 0, aa = 0, digit = 0, alpha = 0;
    const std::string proper_nucleotides = "ACGTU", nucleotides = "ACGTURYWSMKBHDVNX",
                      amino = "ARNDCQEGHILKMFPSTWYV";
    for (const std::string &s : seqs)
        for (char ch : s) {
            if (proper_nucleotides.find(ch) != std::string::npos) proper_nuc++;
            if (nucleotides.find(ch) != std::string::npos) nuc++;
            if (amino.find(ch) != std::string::npos) aa++;
            if (std::isdigit((unsigned char)ch)) digit++;
            if (std::isalpha((unsigned char)ch)) alpha++;
        }
    if (digit != 0 || alpha < 10) return -1;
    if (nuc == alpha) return (double)proper_nuc / alpha > 0.9 ? PLK_SEQ_DNA : -1;
    return (double)aa / alpha > 0.9 ? PLK_SEQ_PROTEIN : -1;
}

constexpr uint8_t kInvalid = 255;

// buildStateMap (1788-1846) for DNA (and codon nucleotides) and protein
void state_map(int type, int su, uint8_t *map) {
    std::memset(map, kInvalid, 256);
    for (char ch : std::string("?-~.!")) map[(unsigned char)ch] = (uint8_t)su;
    if (type == PLK_SEQ_PROTEIN) {
        const char *sym = "ARNDCQEGHILKMFPSTWYV";
        for (int i = 0; i < 20; i++) map[(unsigned char)sym[i]] = (uint8_t)i;
        map[(unsigned char)'X'] = (uint8_t)su;
        map[(unsigned char)'B'] = 20;  // N or D
        map[(unsigned char)'Z'] = 21;  // Q or E
        map[(unsigned char)'J'] = 22;  // I or L
        map[(unsigned char)'*'] = (uint8_t)su;
        map[(unsigned char)'U'] = (uint8_t)su;
        map[(unsigned char)'O'] = (uint8_t)su;
        return;
    }
    // nucleotides: 0-3, ambiguity sets as 3 + bitmask (A=1, C=2, G=4, T=8)
    const char *codes = "ACGTURYWSMKBHDV";
    const int masks[] = {1, 2, 4, 8, 8, 5, 10, 9, 6, 3, 12, 14, 11, 13, 7};
    for (int i = 0; i < 15; i++) {
        const int m = masks[i];
        map[(unsigned char)codes[i]] = (uint8_t)((m & (m - 1)) == 0 ? __builtin_ctz(m) : m + 3);
    }
    map[(unsigned char)'N'] = (uint8_t)su;
    map[(unsigned char)'X'] = (uint8_t)su;
}

// Bitset of states a coded character may be (getAppearance), over nstates <= 64.
uint64_t appearance(int type, int nstates, int su, int s) {
    const uint64_t all = nstates >= 64 ? ~0ull : ((1ull << nstates) - 1);
    if (s == su) return all;
    if (s < nstates) return 1ull << s;
    if (type == PLK_SEQ_DNA) return (uint64_t)(s - 3) & 15;
    if (type == PLK_SEQ_PROTEIN) {
        static const int pair[3][2] = {{2, 3}, {5, 6}, {9, 10}};
        return (1ull << pair[s - 20][0]) | (1ull << pair[s - 20][1]);
    }
    return all;
}

void build_patterns(Align &A, const std::vector<std::string> &seqs, int type) {
    const int nseq = (int)seqs.size();
    const int n = type == PLK_SEQ_DNA ? 4 : type == PLK_SEQ_PROTEIN ? 20 : 61;
    const int su = plk_state_unknown(n, type);
    uint8_t map[256];
    state_map(type == PLK_SEQ_PROTEIN ? PLK_SEQ_PROTEIN : PLK_SEQ_DNA, su, map);
    const int step = type == PLK_SEQ_CODON ? 3 : 1;
    const int len = (int)seqs[0].size();
    if (len % step != 0) throw ParseError{"Number of sites is not multiple of 3"};
    // sense-codon index of each of the 64 codons (-1 for stops)
    int sense[64], ns = 0;
    for (int c = 0; c < 64; c++) sense[c] = kStandardGeneticCode[c] == '*' ? -1 : ns++;
    A.nstates = n;
    A.seq_type = type;
    A.ntaxa = nseq;
    A.nsites = len / step;
    A.site_pattern.assign(A.nsites, -1);
    std::unordered_map<std::string, int> index;
    std::vector<std::string> pats;
    std::string pat(nseq, '\0');
    std::string errors;
    int nerr = 0;
    for (int site = 0; site < len; site += step) {
        for (int q = 0; q < nseq; q++) {
            int s = map[(unsigned char)seqs[q][site]];
            if (type == PLK_SEQ_CODON) {
                const int s2 = map[(unsigned char)seqs[q][site + 1]], s3 = map[(unsigned char)seqs[q][site + 2]];
                if (s < 4 && s2 < 4 && s3 < 4) {
                    const int c = s * 16 + s2 * 4 + s3;
                    s = sense[c] < 0 ? su : sense[c];  // a stop codon is missing data
                } else if (s == kInvalid || s2 == kInvalid || s3 == kInvalid) {
                    s = kInvalid;
                } else {
                    s = su;  // partly ambiguous codon
                }
            }
            if (s == kInvalid) {
                if (nerr < 100)
                    errors += "Sequence " + A.names[q] + " has invalid character " + seqs[q].substr(site, step) +
                              " at site " + std::to_string(site + 1) + "\n";
                nerr++;
            }
            pat[q] = (char)s;
        }
        if (nerr) continue;
        bool gaps = true;
        for (int q = 0; q < nseq && gaps; q++) gaps = (int)(uint8_t)pat[q] == su;
        A.gaps_only += gaps;
        auto it = index.find(pat);
        int id;
        if (it == index.end()) {
            id = (int)pats.size();
            index.emplace(pat, id);
            pats.push_back(pat);
            A.freq.push_back(1.0);
        } else {
            id = it->second;
            A.freq[id] += 1.0;
        }
        A.site_pattern[site / step] = id;
    }
    if (nerr) throw ParseError{errors};
    A.npat = (int)pats.size();
    A.states.resize((size_t)nseq * A.npat);
    A.flags.assign(A.npat, 0);
    for (int p = 0; p < A.npat; p++) {
        uint64_t app = n >= 64 ? ~0ull : ((1ull << n) - 1);
        std::vector<int> count(n, 0);
        for (int q = 0; q < nseq; q++) {
            const int s = (uint8_t)pats[p][q];
            A.states[(size_t)q * A.npat + p] = (uint8_t)s;
            app &= appearance(type, n, su, s);
            if (s < n) count[s]++;
        }
        int twice = 0;
        for (int j = 0; j < n; j++) twice += count[j] >= 2;
        // computeConst (1321-1420): constant when one state fits every character;
        // informative when two states appear at least twice each
        A.flags[p] = (app != 0 ? 1 : 0) | (twice >= 2 ? 2 : 0);
    }
}

int parse(plk_align **out, const char *text, size_t len, int format, int seq_type) {
    if (!out || !text || format < 0 || format > 3 || seq_type < -1 || seq_type > PLK_SEQ_CODON)
        return fail(PLK_ERR_ARG, "plk_align: bad arguments");
    *out = nullptr;
    auto lines = split_lines(text, len);
    if (format == 0) {
        format = 1;
        for (const std::string &l : lines) {
            const size_t k = l.find_first_not_of(" \t");
            if (k == std::string::npos) continue;
            format = l[k] == '>' ? 3 : 1;
            break;
        }
    }
    auto *a = new plk_align();
    try {
        std::vector<std::string> seqs;
        int nsite;
        if (format == 1) nsite = read_phylip(lines, a->names, seqs);
        else if (format == 2) nsite = read_phylip_sequential(lines, a->names, seqs);
        else nsite = read_fasta(lines, a->names, seqs);
        // buildPattern's checks (2219-2265)
        std::unordered_set<std::string> seen;
        std::string err;
        for (size_t i = 0; i < a->names.size(); i++) {
            if (a->names[i].empty()) err += "Sequence number " + std::to_string(i + 1) + " has no names\n";
            if (!seen.insert(a->names[i]).second) err += "The sequence name " + a->names[i] + " is duplicated\n";
        }
        if (!err.empty()) throw ParseError{err};
        for (size_t i = 0; i < seqs.size(); i++)
            if ((int)seqs[i].size() != nsite)
                err += "Sequence " + a->names[i] + " contains " + ((int)seqs[i].size() < nsite ? "not enough" : "too many") +
                       " characters (" + std::to_string(seqs[i].size()) + ")\n";
        if (!err.empty()) throw ParseError{err};
        if (seqs.size() < 3) throw ParseError{"There must be at least 3 sequences"};
        if (seqs[0].empty()) throw ParseError{"No alignment columns"};
        int type = seq_type;
        if (type < 0) {
            type = detect_type(seqs);
            if (type < 0) throw ParseError{"Unknown sequence type."};
        }
        build_patterns(*a, seqs, type);
    } catch (const ParseError &e) {
        delete a;
        return fail(PLK_ERR_ARG, e.msg);
    }
    *out = a;
    return PLK_OK;
}

}  // namespace
}  // namespace plk

using namespace plk;

extern "C" int plk_align_parse(plk_align **out, const char *text, size_t len, int format, int seq_type) {
    clear_error();
    return plk::parse(out, text, len, format, seq_type);
}

extern "C" int plk_align_read(plk_align **out, const char *path, int format, int seq_type) {
    clear_error();
    if (!out || !path) return fail(PLK_ERR_ARG, "plk_align_read: bad arguments");
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(PLK_ERR_ARG, std::string("plk_align_read: cannot open ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    return plk::parse(out, text.data(), text.size(), format, seq_type);
}

extern "C" int plk_align_destroy(plk_align *a) {
    delete a;
    return PLK_OK;
}

extern "C" int plk_align_dims(const plk_align *a, int32_t *dims) {
    if (!a || !dims) return fail(PLK_ERR_ARG, "plk_align_dims: null");
    dims[0] = a->ntaxa;
    dims[1] = a->nsites;
    dims[2] = a->npat;
    dims[3] = a->nstates;
    dims[4] = a->seq_type;
    dims[5] = a->gaps_only;
    return PLK_OK;
}

extern "C" int plk_align_name(const plk_align *a, int taxon, char *buf, int buflen) {
    if (!a || taxon < 0 || taxon >= a->ntaxa || !buf || buflen < 1) return fail(PLK_ERR_ARG, "plk_align_name: bad arguments");
    std::snprintf(buf, (size_t)buflen, "%s", a->names[taxon].c_str());
    return PLK_OK;
}

extern "C" int plk_align_patterns(const plk_align *a, uint8_t *states, double *freq, int32_t *site_pattern,
                                  int32_t *flags) {
    if (!a) return fail(PLK_ERR_ARG, "plk_align_patterns: null");
    if (states) std::memcpy(states, a->states.data(), a->states.size());
    if (freq) std::memcpy(freq, a->freq.data(), sizeof(double) * a->freq.size());
    if (site_pattern) std::memcpy(site_pattern, a->site_pattern.data(), sizeof(int32_t) * a->site_pattern.size());
    if (flags) std::memcpy(flags, a->flags.data(), sizeof(int32_t) * a->flags.size());
    return PLK_OK;
}

// mc_policy.cpp — the reference's Monte Carlo playout policy (see
// mc_policy.h). The decision steps follow the reference's move generators
// and readers (FastState.cpp:75-255, FastBoard.cpp:191-343, 725-990,
// 1179-1471, 1486-1523, 1686-1773, 2027-2643, Matcher.cpp:14-112) because
// the labels are defined by them; the tables are the reference's data
// (leela_amd/data/mc_policy.bin).
#include "mc_policy.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "playout.h"

namespace lz {
namespace mc {

namespace {

constexpr int kBlack = go::BLACK, kWhite = go::WHITE, kEmpty = go::EMPTY,
              kInval = go::OFFBOARD;
constexpr float kBound = 32.0f;   // cfg_bound (GTP.cpp:85)
constexpr int kRandomLoops = 4;   // cfg_random_loops (GTP.cpp:100)

uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

// The perfect hash of a pattern (Matcher::PatHashG / PatHashV).
uint32_t pattern_hash(const std::vector<int32_t> &g, uint32_t p) {
    uint32_t q = p;
    q ^= q >> 3;
    const uint32_t d = (uint32_t)g[q % Tables::kGSize];
    p ^= rotl32(p, 11) ^ ((p + d) >> 6);
    return p % Tables::kVSize;
}

// The smallest of a 3x3 pattern's eight symmetric encodings, colours
// swapped first when `invert` (FastBoard::get_pattern3_augment_spec):
// s[0..7] the squares NW, N, NE, W, E, SW, S, SE, a[0..3] the liberty
// codes of N, W, E, S.
uint32_t canonical(const int (&s0)[8], const int (&a)[4], bool invert) {
    int s[8];
    for (int i = 0; i < 8; i++)
        s[i] = invert && s0[i] < 2 ? 1 - s0[i] : s0[i];
    // Each symmetry: the order in which the squares and the liberty codes
    // are read.
    static const int sq_order[8][8] = {
        {0, 1, 2, 3, 4, 5, 6, 7}, {5, 3, 0, 6, 1, 7, 4, 2},
        {7, 6, 5, 4, 3, 2, 1, 0}, {2, 4, 7, 1, 6, 0, 3, 5},
        {0, 3, 5, 1, 6, 2, 4, 7}, {2, 1, 0, 4, 3, 7, 6, 5},
        {7, 4, 2, 6, 1, 5, 3, 0}, {5, 6, 7, 3, 4, 0, 1, 2}};
    static const int lib_order[8][4] = {
        {0, 1, 2, 3}, {1, 3, 0, 2}, {3, 2, 1, 0}, {2, 0, 3, 1},
        {1, 0, 3, 2}, {0, 2, 1, 3}, {2, 3, 0, 1}, {3, 1, 2, 0}};
    uint32_t best = 0xffffffffu;
    for (int k = 0; k < 8; k++) {
        uint32_t v = 0;
        for (int i = 0; i < 8; i++) v = (v << 2) | (uint32_t)s[sq_order[k][i]];
        for (int i = 0; i < 4; i++)
            v |= (uint32_t)a[lib_order[k][i]] << (22 - 2 * i);
        best = std::min(best, v);
    }
    return best;
}

std::string library_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&library_dir), &info) &&
        info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t slash = p.rfind('/');
        return slash == std::string::npos ? std::string(".")
                                          : p.substr(0, slash);
    }
    return ".";
}

template <typename T>
bool read_array(FILE *f, std::vector<T> &v, size_t n) {
    v.resize(n);
    return std::fread(v.data(), sizeof(T), n, f) == n;
}

bool load_tables(const std::string &path, Tables &t, std::string *err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        *err = "cannot open " + path;
        return false;
    }
    char magic[8];
    int32_t n[5];
    std::vector<int32_t> patterns, live;
    std::vector<float> fw, fw_sl, pw, pw_sl;
    bool ok = std::fread(magic, 1, 8, f) == 8 &&
              std::memcmp(magic, "LZMCPOL1", 8) == 0 &&
              std::fread(n, sizeof n, 1, f) == 1 && n[1] == Tables::kGSize &&
              n[3] == kNumFeatures && n[4] == n[2] &&
              read_array(f, patterns, n[0]) && read_array(f, t.g, n[1]) &&
              read_array(f, live, n[2]) && read_array(f, fw, n[3]) &&
              read_array(f, fw_sl, n[3]) && read_array(f, pw, n[4]) &&
              read_array(f, pw_sl, n[4]);
    std::fclose(f);
    if (!ok) {
        *err = "bad playout policy table file " + path;
        return false;
    }
    // Weights times their SL factors (cfg_mc_softmax = 1: the power is the
    // identity).
    for (int i = 0; i < kNumFeatures; i++)
        t.feature_weight[i] = fw[i] * fw_sl[i];
    t.pattern_weight.resize(pw.size());
    for (size_t i = 0; i < pw.size(); i++)
        t.pattern_weight[i] = pw[i] * pw_sl[i];
    // Matcher::Matcher: every valid pattern's hash slot holds the index of
    // its canonical form among the weighted patterns, per colour to move.
    std::unordered_map<uint32_t, int> index_of;
    for (size_t j = 0; j < live.size(); j++)
        index_of.emplace((uint32_t)live[j], (int)j);
    t.index[0].assign(Tables::kVSize, 0);
    t.index[1].assign(Tables::kVSize, 0);
    for (const int32_t pat : patterns) {
        const uint32_t p = (uint32_t)pat;
        int s[8], a[4];
        for (int i = 0; i < 8; i++) s[i] = (p >> (14 - 2 * i)) & 3;
        for (int i = 0; i < 4; i++) a[i] = (p >> (22 - 2 * i)) & 3;
        const uint32_t h = pattern_hash(t.g, p);
        for (int c = 0; c < 2; c++) {
            auto it = index_of.find(canonical(s, a, c == 1));
            if (it != index_of.end()) t.index[c][h] = (uint16_t)it->second;
        }
    }
    return true;
}

}  // namespace

int Tables::matches(int c, uint32_t pattern) const {
    return index[c][pattern_hash(g, pattern)];
}

const Tables *tables(const char *path, std::string *err) {
    static std::mutex mu;
    static Tables *loaded = nullptr;
    static std::string load_err;
    std::lock_guard<std::mutex> g(mu);
    if (path) {  // an explicit file: load it (kept for the process)
        auto *t = new Tables();
        if (!load_tables(path, *t, err)) {
            delete t;
            return nullptr;
        }
        loaded = t;
        return t;
    }
    if (!loaded && load_err.empty()) {
        const char *env = std::getenv("LZ_MC_POLICY");
        const std::string file =
            env ? std::string(env) : library_dir() + "/data/mc_policy.bin";
        auto *t = new Tables();
        if (load_tables(file, *t, &load_err))
            loaded = t;
        else
            delete t;
    }
    if (!loaded) *err = load_err;
    return loaded;
}

// ---- PlayBoard ---------------------------------------------------------

void PlayBoard::reset(int size) {
    size_ = size;
    const int W = size + 2;
    dirs_[0] = -W;
    dirs_[1] = +1;
    dirs_[2] = +W;
    dirs_[3] = -1;
    const int ex[8] = {-W - 1, -W, -W + 1, -1, +1, W - 1, W, W + 1};
    for (int i = 0; i < 8; i++) extradirs_[i] = ex[i];
    for (int i = 0; i < kMaxSq; i++) {
        sq_[i] = kInval;
        nbr_[i] = 0;
        parent_[i] = kNone;
    }
    empty_cnt_ = 0;
    // Empty points column by column, with the border counting as both
    // colours in the neighbour counts.
    for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) {
            const int v = vertex(i, j);
            sq_[v] = kEmpty;
            empty_idx_[v] = empty_cnt_;
            empty_[empty_cnt_++] = v;
            int n = 0;
            n += (i == 0 || i == size - 1) ? (1 | (1 << 4) | (1 << 8))
                                           : (2 << 8);
            n += (j == 0 || j == size - 1) ? (1 | (1 << 4) | (1 << 8))
                                           : (2 << 8);
            nbr_[v] = (uint16_t)n;
        }
    parent_[kNone] = kNone;
    libs_[kNone] = 16384;  // large: borders and empties never count low
    next_[kNone] = kNone;
    stones_[kNone] = 0;
    total_[0] = total_[1] = 0;
    critical_.clear();
}

void PlayBoard::add_neighbour(int v, int color) {
    int pars[4], cnt = 0;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        nbr_[ai] = (uint16_t)(nbr_[ai] + (1 << (4 * color)) - (1 << 8));
        const int p = parent_[ai];
        bool found = false;
        for (int i = 0; i < cnt; i++) found |= pars[i] == p;
        if (!found) {
            libs_[p]--;
            pars[cnt++] = p;
        }
    }
}

void PlayBoard::remove_neighbour(int v, int color) {
    int pars[4], cnt = 0;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        nbr_[ai] = (uint16_t)(nbr_[ai] + (1 << 8) - (1 << (4 * color)));
        const int p = parent_[ai];
        bool found = false;
        for (int i = 0; i < cnt; i++) found |= pars[i] == p;
        if (!found) {
            libs_[p]++;
            pars[cnt++] = p;
        }
    }
}

int PlayBoard::remove_string(int v) {
    const int color = sq_[v];
    int pos = v, removed = 0;
    do {
        sq_[pos] = kEmpty;
        parent_[pos] = kNone;
        total_[color]--;
        remove_neighbour(pos, color);
        empty_idx_[pos] = empty_cnt_;
        empty_[empty_cnt_++] = pos;
        removed++;
        pos = next_[pos];
    } while (pos != v);
    return removed;
}

void PlayBoard::merge_strings(int ip, int aip) {
    stones_[ip] += stones_[aip];
    int pos = aip;
    do {
        // Liberties of the joining stones that ip's string lacks.
        for (int k = 0; k < 4; k++) {
            const int ai = pos + dirs_[k];
            if (sq_[ai] != kEmpty) continue;
            bool shared = false;
            for (int kk = 0; kk < 4; kk++)
                shared |= parent_[ai + dirs_[kk]] == ip;
            if (!shared) libs_[ip]++;
        }
        parent_[pos] = ip;
        pos = next_[pos];
    } while (pos != aip);
    std::swap(next_[aip], next_[ip]);
}

void PlayBoard::take_empty(int v) {
    const int last = empty_[--empty_cnt_];
    empty_idx_[last] = empty_idx_[v];
    empty_[empty_idx_[v]] = last;
}

int PlayBoard::place_eye(int color, int v) {
    sq_[v] = (uint8_t)color;
    next_[v] = v;
    parent_[v] = v;
    libs_[v] = 0;
    stones_[v] = 1;
    total_[color]++;
    add_neighbour(v, color);
    int captured_sq = -1, captured = 0;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (libs_at(ai) <= 0) {
            captured += remove_string(ai);
            captured_sq = ai;
        }
    }
    take_empty(v);
    return captured == 1 ? captured_sq : -1;
}

int PlayBoard::play_fast(int color, int v, bool *capture) {
    if (nbr_[v] & (4 << (4 * (1 - color)))) {  // into an opponent eye
        *capture = true;
        return place_eye(color, v);
    }
    sq_[v] = (uint8_t)color;
    next_[v] = v;
    parent_[v] = v;
    libs_[v] = count(kEmpty, v);
    stones_[v] = 1;
    total_[color]++;
    add_neighbour(v, color);
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (sq_[ai] > kWhite) continue;
        if (sq_[ai] == 1 - color) {
            if (libs_at(ai) <= 0) {
                *capture = true;
                remove_string(ai);
            }
        } else {
            const int ip = parent_[v], aip = parent_[ai];
            if (ip != aip) {
                if (stones_[ip] >= stones_[aip])
                    merge_strings(ip, aip);
                else
                    merge_strings(aip, ip);
            }
        }
    }
    take_empty(v);
    if (libs_at(v) == 0) remove_string(v);  // suicide
    return -1;
}

void PlayBoard::setup(const go::Board &b) {
    reset(b.side());
    for (int y = 0; y < size_; y++)
        for (int x = 0; x < size_; x++) {
            const int v = vertex(x, y);
            const int c = b.at(b.vertex(x, y));
            if (c != kBlack && c != kWhite) continue;
            // A game move (FullBoard::update_board) that captures nothing:
            // in a legal position every string keeps a liberty.
            sq_[v] = (uint8_t)c;
            next_[v] = v;
            parent_[v] = v;
            libs_[v] = count(kEmpty, v);
            stones_[v] = 1;
            total_[c]++;
            add_neighbour(v, c);
            for (int k = 0; k < 4; k++) {
                const int ai = v + dirs_[k];
                if (sq_[ai] != c) continue;
                const int ip = parent_[v], aip = parent_[ai];
                if (ip != aip) {
                    if (stones_[ip] >= stones_[aip])
                        merge_strings(ip, aip);
                    else
                        merge_strings(aip, ip);
                }
            }
            take_empty(v);
        }
}

float PlayBoard::final_mc_score(float komi) const {
    int bsc = total_[kBlack], wsc = total_[kWhite];
    for (int i = 0; i < empty_cnt_; i++) {
        const int v = empty_[i];
        if (count(kWhite, v) == 4)
            wsc++;
        else if (count(kBlack, v) == 4)
            bsc++;
    }
    return (float)bsc - ((float)wsc + komi);
}

bool PlayBoard::is_eye(int color, int v) const {
    if (!(nbr_[v] & (4 << (4 * color)))) return false;
    const int W = size_ + 2;
    int cnt[4] = {0, 0, 0, 0};
    cnt[sq_[v - 1 - W]]++;
    cnt[sq_[v + 1 - W]]++;
    cnt[sq_[v - 1 + W]]++;
    cnt[sq_[v + 1 + W]]++;
    if (cnt[kInval] == 0) return cnt[1 - color] <= 1;
    return cnt[1 - color] == 0;
}

bool PlayBoard::fast_ss_suicide(int color, int v) const {
    if (!(nbr_[v] & (4 << (4 * (1 - color))))) return false;
    const int W = size_ + 2;
    if (libs_at(v - 1) <= 1) return false;
    if (libs_at(v + 1) <= 1) return false;
    if (libs_at(v + W) <= 1) return false;
    if (libs_at(v - W) <= 1) return false;
    return true;
}

bool PlayBoard::is_suicide(int v, int color) {
    if (count(kEmpty, v)) return false;
    bool connecting = false;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        const int libs = libs_at(ai);
        if (sq_[ai] == color) {
            if (libs > 1) return false;  // joins a string that lives
            connecting = true;
        } else if (libs <= 1) {
            return false;  // captures
        }
    }
    add_neighbour(v, color);
    bool opps_live = true, ours_die = true;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        const int libs = libs_at(ai);
        if (libs == 0 && sq_[ai] != color)
            opps_live = false;
        else if (libs != 0 && sq_[ai] == color)
            ours_die = false;
    }
    remove_neighbour(v, color);
    return connecting ? (opps_live && ours_die) : opps_live;
}

template <int N>
void PlayBoard::add_string_liberties(int v, std::array<int, N> &libs,
                                     size_t &cnt) const {
    int pos = v;
    do {
        if (count(kEmpty, pos)) {
            for (int k = 0; k < 4; k++) {
                const int ai = pos + dirs_[k];
                if (sq_[ai] != kEmpty) continue;
                bool found = false;
                for (size_t i = 0; i < cnt; i++)
                    if (libs[i] == ai) {
                        found = true;
                        break;
                    }
                if (!found) {
                    libs[cnt++] = ai;
                    if (cnt >= N) return;
                }
            }
        }
        pos = next_[pos];
    } while (pos != v);
}

bool PlayBoard::kill_or_connect(int color, int v) const {
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        const int s = sq_[ai], libs = libs_at(ai);
        if ((libs <= 1 && s == 1 - color) || (libs >= 3 && s == color))
            return true;
    }
    return false;
}

bool PlayBoard::self_atari(int color, int v) const {
    if (count(kEmpty, v) >= 2) return false;
    if (kill_or_connect(color, v)) return false;
    if (count(color, v) == 0) return true;
    std::array<int, 3> libs;
    size_t cnt = 0;
    libs[cnt++] = v;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (sq_[ai] == kEmpty) {
            bool found = false;
            for (size_t i = 0; i < cnt; i++)
                if (libs[i] == ai) {
                    found = true;
                    break;
                }
            if (!found) {
                if (cnt > 1) return false;
                libs[cnt++] = ai;
            }
        } else if (sq_[ai] == color) {
            if (libs_at(ai) > 1) {
                add_string_liberties<3>(ai, libs, cnt);
                if (cnt > 2) return false;
            }
        }
    }
    return true;
}

int PlayBoard::enemy_atari_size(int color, int v) const {
    int size = 0;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (sq_[ai] == 1 - color && libs_at(ai) <= 2)
            size = std::max(size, string_size(ai));
    }
    return size;
}

uint32_t PlayBoard::pattern_augment(int v) const {
    const int W = size_ + 2;
    const int nb[8] = {v - W - 1, v - W, v - W + 1, v - 1,
                       v + 1,     v + W - 1, v + W, v + W + 1};
    uint32_t res = 0;
    for (int i = 0; i < 8; i++) res = (res << 2) | sq_[nb[i]];
    // Liberty codes of the N, W, E and S strings: min(liberties - 1, 3)
    // (-1 for a string without liberties, as the reference's arithmetic
    // gives it: all high bits set).
    const int side4[4] = {v - W, v - 1, v + 1, v + W};
    for (int i = 0; i < 4; i++) {
        const int at = std::min(libs_at(side4[i]) - 1, 3);
        res |= (uint32_t)at << (22 - 2 * i);
    }
    return res;
}

int PlayBoard::in_atari(int v) const {
    if (libs_at(v) > 1) return 0;
    int pos = v;
    do {
        if (count(kEmpty, pos))
            for (int k = 0; k < 4; k++) {
                const int ai = pos + dirs_[k];
                if (sq_[ai] == kEmpty) return ai;
            }
        pos = next_[pos];
    } while (pos != v);
    return 0;
}

namespace {

Move capture_move(int v, int size, const Tables &t) {
    return Move(v, size == 1 ? kCapture1 : size == 2 ? kCapture2 : kCapture3P,
                t);
}

}  // namespace

bool PlayBoard::kill_neighbours(int v, int komove, std::vector<Move> &m,
                                const Tables &t) {
    bool result = false;
    const int kcolor = 1 - sq_[v];
    std::array<int, 4> seen;
    int nseen = 0;
    int pos = v;
    do {
        for (int k = 0; k < 4; k++) {
            const int ai = pos + dirs_[k];
            if (sq_[ai] != kcolor) continue;
            const int par = parent_[ai];
            if (libs_[par] <= 1 && nseen < 4) {
                bool found = false;
                for (int i = 0; i < nseen; i++) found |= seen[i] == par;
                if (!found) {
                    const int atari = in_atari(ai);
                    result = true;
                    if (atari != komove) {
                        m.push_back(capture_move(atari, stones_[par], t));
                        const int sav = string_size(v);
                        m.back().add(sav == 1   ? kSaving1
                                     : sav == 2 ? kSaving2
                                                : kSaving3P,
                                     t);
                        m.back().add(kSavingKill, t);
                    }
                    seen[nseen++] = par;
                }
            }
        }
        pos = next_[pos];
    } while (pos != v);
    return result;
}

void PlayBoard::save_critical_neighbours(int color, int v, int komove,
                                         std::vector<Move> &m,
                                         const Tables &t) {
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (sq_[ai] != color) continue;
        const int par = parent_[ai];
        if (libs_[par] > 1) continue;
        const int atari = in_atari(ai);
        if (atari != komove) {
            const int size = stones_[par];
            Move sm(atari, size == 1 ? kSaving1 : size == 2 ? kSaving2
                                                            : kSaving3P,
                    t);
            const int libs = count(kEmpty, atari);
            if (libs == 2) sm.add(kSaving2Lib, t);
            if (libs == 3) sm.add(kSaving3Lib, t);
            m.push_back(sm);
        }
        const bool sa = self_atari(color, atari);
        const bool kill = kill_neighbours(ai, komove, m, t);
        if (!kill && (sa || atari == komove)) critical_.push_back(atari);
    }
}

void PlayBoard::try_capture(int color, int v, std::vector<Move> &m,
                            const Tables &t) {
    if (sq_[v] != kEmpty || !count(1 - color, v)) return;
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (sq_[ai] == 1 - color && libs_at(ai) <= 1) {
            m.push_back(capture_move(v, string_size(ai), t));
            return;
        }
    }
}

void PlayBoard::add_global_captures(int color, int komove,
                                    std::vector<Move> &m, const Tables &t) {
    for (const int v : critical_)
        if (v != komove) try_capture(color, v, m, t);
    critical_.clear();
}

template <int N>
bool PlayBoard::saveable_string(int parent, std::array<int, N> &libs,
                                size_t &cnt) const {
    add_string_liberties<N>(parent, libs, cnt);
    for (size_t i = 0; i < cnt; i++) {
        const int lib = libs[i];
        for (int k = 0; k < 4; k++) {
            const int ai = lib + dirs_[k];
            if (sq_[ai] == kEmpty) {
                if (std::find(libs.begin(), libs.begin() + cnt, ai) ==
                    libs.begin() + cnt)
                    return true;  // a new liberty next to one
            } else if (sq_[ai] == sq_[parent]) {
                const int pp = parent_[ai];
                if (pp == parent) continue;
                if (libs_[pp] > (int)cnt) return true;
                if (libs_[pp] >= 2) {
                    // Connecting: do the joined liberties outnumber ours?
                    std::array<int, N> other;
                    size_t ocnt = 0;
                    add_string_liberties<N>(pp, other, ocnt);
                    std::vector<int> all(other.begin(), other.begin() + ocnt);
                    all.insert(all.end(), libs.begin(), libs.begin() + cnt);
                    std::sort(all.begin(), all.end());
                    all.erase(std::unique(all.begin(), all.end()), all.end());
                    if (all.size() > cnt) return true;
                }
            }
        }
    }
    return false;
}

void PlayBoard::add_semeai_moves(int color, int last, int komove,
                                 std::vector<Move> &m, const Tables &t) {
    constexpr int kMaxLibs = 3;
    std::vector<int> own;
    for (int k = 0; k < 4; k++) {
        const int ai = last + dirs_[k];
        if (sq_[ai] == color && libs_at(ai) <= kMaxLibs)
            own.push_back(parent_[ai]);
    }
    for (const int og : own) {
        if (libs_[og] <= 1) continue;  // strings in atari are handled apart
        std::array<int, kMaxLibs> olibs;
        size_t ocnt = 0;
        if (saveable_string<kMaxLibs>(og, olibs, ocnt)) continue;
        // Enemy strings around with no more liberties than ours.
        std::vector<int> enemies;
        int pos = og;
        do {
            for (int k = 0; k < 4; k++) {
                const int ai = pos + dirs_[k];
                if (sq_[ai] == 1 - sq_[og] && libs_at(ai) <= (int)ocnt)
                    enemies.push_back(parent_[ai]);
            }
            pos = next_[pos];
        } while (pos != og);
        for (const int eg : enemies) {
            if (libs_[eg] <= 1) continue;  // captures are handled apart
            std::array<int, kMaxLibs> elibs;
            size_t ecnt = 0;
            if (saveable_string<kMaxLibs>(eg, elibs, ecnt)) continue;
            for (size_t i = 0; i < ecnt; i++) {
                const int lib = elibs[i];
                if (lib == komove) continue;
                const bool ours = std::find(olibs.begin(),
                                            olibs.begin() + ocnt, lib) !=
                                  olibs.begin() + ocnt;
                if (!ours && !fast_ss_suicide(color, lib))
                    m.emplace_back(lib, ocnt <= 2 ? kSemeai2 : kSemeai3, t);
            }
        }
    }
}

void PlayBoard::check_nakade(int color, int v, int komove,
                             std::vector<Move> &m, const Tables &t) {
    std::array<int, 6> nakade, empty_counts{};
    std::array<int, 5> nbr_to_coord{};
    const int nbrs = count(kEmpty, v);
    if (nbrs == 0) return;
    size_t cnt = 0;
    nbr_to_coord[nbrs] = v;
    empty_counts[nbrs]++;
    nakade[cnt++] = v;
    size_t idx = 0;
    do {
        const int nv = nakade[idx];
        for (int k = 0; k < 4; k++) {
            const int ai = nv + dirs_[k];
            if (sq_[ai] == 1 - color) return;  // not surrounded
            if (sq_[ai] != kEmpty) continue;
            bool found = false;
            for (size_t j = 0; j < cnt; j++) found |= nakade[j] == ai;
            if (found) continue;
            if (cnt >= nakade.size()) return;  // too much eye space
            const int n = count(kEmpty, ai);
            nbr_to_coord[n] = ai;
            empty_counts[n]++;
            nakade[cnt++] = ai;
        }
    } while (++idx < cnt);
    auto corner = [&](int p) {
        const int W = size_ + 2, x = p % W - 1, y = p / W - 1;
        return (x == 0 || x == size_ - 1) && (y == 0 || y == size_ - 1);
    };
    auto add = [&](int p) {
        if (p != komove) m.emplace_back(p, kNakade, t);
    };
    // Killable eye shapes (senseis.xmp.net/?KillableEyeShapes).
    if (cnt <= 2) return;
    if (cnt == 3) {
        add(nbr_to_coord[2]);
    } else if (cnt == 4) {
        if (empty_counts[3] == 1) {  // pyramid four
            add(nbr_to_coord[3]);
        } else if (empty_counts[2] == 2 && empty_counts[1] == 2) {
            // bent four in the corner
            int crit = 0;
            bool bent4 = false;
            for (size_t j = 0; j < cnt; j++) {
                const int p = nakade[j];
                if (corner(p)) {
                    if (count(kEmpty, p) == 2) bent4 = true;
                } else if (count(kEmpty, p) == 2) {
                    crit = p;
                }
            }
            if (bent4) add(crit);
        }
    } else if (cnt == 5) {
        if (empty_counts[1] == 1 && empty_counts[3] == 1)  // bulky five
            add(nbr_to_coord[3]);
        else if (empty_counts[4] == 1)  // crossed five
            add(nbr_to_coord[4]);
    } else if (cnt == 6) {
        if (empty_counts[1] == 2 && empty_counts[4] == 1) {  // rabbity six
            add(nbr_to_coord[4]);
        } else if (empty_counts[2] == 4 && empty_counts[3] == 2) {
            // rectangular six in the corner
            int crit = 0;
            bool rect6 = false;
            for (size_t j = 0; j < cnt; j++) {
                const int p = nakade[j];
                if (corner(p))
                    rect6 = true;
                else if (count(kEmpty, p) == 3 && count(color, p) == 1 &&
                         count(1 - color, p) == 0)
                    crit = p;
            }
            if (rect6) add(crit);
        }
    }
}

void PlayBoard::add_near_nakade_moves(int color, int v, int komove,
                                      std::vector<Move> &m,
                                      const Tables &t) {
    for (int k = 0; k < 4; k++) {
        const int ai = v + dirs_[k];
        if (sq_[ai] == kEmpty) check_nakade(1 - color, ai, komove, m, t);
    }
}

void PlayBoard::add_pattern_moves(int color, int v, int komove,
                                  std::vector<Move> &m, const Tables &t) {
    for (int i = 0; i < 8; i++) {
        const int p = v + extradirs_[i];
        if (sq_[p] == kEmpty && p != komove && !fast_ss_suicide(color, p))
            m.emplace_back(p, kPattern, t);
    }
}

// ---- PlayState -----------------------------------------------------------

void PlayState::setup(const go::GameState &st) {
    board.setup(st.board);
    to_move = st.to_move();
    komove = st.komove;
    last_move = st.last_move;
    last_was_capture = st.last_was_capture;
    passes = st.passes;
    movenum = st.movenum;
    komi = st.komi;
}

int PlayState::play_fast(int v) {
    bool capture = false;
    if (v == go::kPass) {
        if (++passes > 4) passes = 4;
    } else {
        komove = board.play_fast(to_move, v, &capture);
        passes = 0;
    }
    last_was_capture = capture;
    last_move = v;
    to_move = 1 - to_move;
    movenum++;
    return v;
}

bool PlayState::try_move(int color, int v) const {
    return v != komove && board.no_eye_fill(v, to_move) &&
           !board.fast_ss_suicide(color, v);
}

int PlayState::walk_empty_list(int color, Rng &rng) {
    const int dir = (int)rng.randfix<2>();
    const int n = board.empty_count();
    const int start = (int)rng.randuint16((uint16_t)n);
    if (dir == 0) {
        for (int i = start; i < n; i++)
            if (try_move(color, board.empty_at(i))) return board.empty_at(i);
        for (int i = 0; i < start; i++)
            if (try_move(color, board.empty_at(i))) return board.empty_at(i);
    } else {
        for (int i = start; i >= 0; i--)
            if (try_move(color, board.empty_at(i))) return board.empty_at(i);
        for (int i = n - 1; i > start; i--)
            if (try_move(color, board.empty_at(i))) return board.empty_at(i);
    }
    return go::kPass;
}

void PlayState::flag_move(Move &m, int color, const Tables &t) {
    m.score *= t.pattern_weight[t.matches(color, board.pattern_augment(m.sq))];
    if (!board.self_atari(color, m.sq)) return;
    if (board.is_suicide(m.sq, color)) {
        m.add(kSuicide, t);
        return;
    }
    if (board.enemy_atari_size(1 - color, m.sq) >= 6) {
        // a self-atari of six stones leaves a living eye space of seven
        m.add(kTooBigSA, t);
    } else {
        const int dying = board.enemy_atari_size(color, m.sq);
        m.add(dying >= 5 ? kForceBigSA : dying ? kForceSA : kSA, t);
    }
    if (m.has(kPattern)) m.add(kPatternSA, t);
    if (m.has(kSaving1) || m.has(kSaving2) || m.has(kSaving3P))
        m.add(kSavingSA, t);
}

int PlayState::play_random_move(Rng &rng, const Tables &t) {
    const int color = to_move;
    const int W = board.size() + 2;
    moves.clear();
    // Local and tactical moves around the opponent's last stone.
    if (last_move > 0 && last_move < W * W && board.at(last_move) == 1 - color) {
        board.add_global_captures(color, komove, moves, t);
        board.save_critical_neighbours(color, last_move, komove, moves, t);
        if (moves.empty())
            board.add_semeai_moves(color, last_move, komove, moves, t);
        if (last_was_capture || (0.3f > rng.randflt()))
            board.add_near_nakade_moves(color, last_move, komove, moves, t);
        board.add_pattern_moves(color, last_move, komove, moves, t);
    }
    float cumul = 0.0f, best = 0.0f;
    std::vector<float> cum;
    cum.reserve(moves.size() + kRandomLoops + 1);
    for (auto &m : moves) {
        flag_move(m, color, t);
        cumul += m.score;
        cum.push_back(cumul);
        best = std::max(best, m.score);
    }
    // No good move yet: random moves from the empty list, and a pass.
    if (moves.empty() || best < kBound) {
        const size_t scored = moves.size();
        for (int loops = 0; loops < kRandomLoops; loops++) {
            const int v = walk_empty_list(color, rng);
            if (v == go::kPass) break;
            moves.emplace_back(v, kRandom, t);
        }
        moves.emplace_back(go::kPass, kPass, t);
        for (size_t i = scored; i < moves.size(); i++) {
            if (moves[i].sq != go::kPass) flag_move(moves[i], color, t);
            cumul += moves[i].score;
            cum.push_back(cumul);
        }
    }
    const float index = rng.randflt() * cumul;
    for (size_t i = 0; i < moves.size(); i++)
        if (index <= cum[i]) return play_fast(moves[i].sq);
    return play_fast(go::kPass);
}

float pattern_playout(const go::GameState &st, Rng &rng, const Tables &t) {
    PlayState ps;
    ps.setup(st);
    const int S = st.board.side();
    const int length = 2 * S * S;
    while (ps.passes < 4 && ps.movenum < length) ps.play_random_move(rng, t);
    return ps.board.final_mc_score(ps.komi);
}

float pattern_winrate(const go::GameState &st, int n, uint32_t seed,
                      const Tables &t) {
    float bwins = 0.0f;
    for (int i = 0; i < n; i++) {
        Rng rng(go::playout_seed(seed, (uint32_t)i));
        rng.next();
        rng.next();
        const float score = pattern_playout(st, rng, t);
        if (score == 0.0f)
            bwins += 0.5f;
        else if (score > 0.0f)
            bwins += 1.0f;
    }
    float w = n > 0 ? bwins / (float)n : 0.5f;
    if (st.to_move() != go::BLACK) w = 1.0f - w;
    return w;
}

}  // namespace mc
}  // namespace lz

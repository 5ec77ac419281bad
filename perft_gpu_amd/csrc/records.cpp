// records.cpp -- perft of every position of a FEN file, with resume.
//
// The reference's records mode (processPerftRecords, launcher.h:1578-1695)
// reads one FEN per line whose last number is an occurrence count, writes
// "<line> <perft> <perft * count>" per line to <file>.op, and on a restart
// skips as many input lines as the .op file already holds (:1627-1631).
// Here the pending lines go to the device in chunks of one batched launch
// each (pg_perft_batch), and every finished chunk is appended and flushed
// before the next starts, so an interrupted run resumes after the last
// complete line (a line left half written is dropped and answered again).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/perft_gpu.h"

namespace {

void strip_newline(std::string &s) {
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
}

// the last space-separated number of a line (the reference's atoi of the
// last token); 0 if the line has none
uint64_t occurrence(const std::string &line) {
    const size_t sp = line.find_last_of(' ');
    const char *tok = line.c_str() + (sp == std::string::npos ? 0 : sp + 1);
    return strtoull(tok, nullptr, 10);
}

bool read_lines(const char *path, std::vector<std::string> &out) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    char buf[4096];
    std::string cur;
    while (fgets(buf, sizeof(buf), f)) {
        cur += buf;
        if (!cur.empty() && cur.back() == '\n') {
            out.push_back(cur);
            cur.clear();
        }
    }
    if (!cur.empty()) out.push_back(cur);  // last line without a newline
    fclose(f);
    return true;
}

}  // namespace

extern "C" int pg_perft_records(pg_ctx *ctx, pg_tt *tt, const char *in_path, uint32_t depth, int slider,
                                uint32_t chunk, uint32_t *done, uint32_t *skipped) {
    if (!ctx || !in_path || depth > 15) return PG_EINVAL;
    if (chunk == 0) chunk = 1024;
    std::vector<std::string> lines;
    if (!read_lines(in_path, lines)) return PG_EINVAL;
    const std::string op = std::string(in_path) + ".op";
    // lines already answered: the complete lines of the output; a line cut
    // short by an interrupted run is dropped and answered again
    std::vector<std::string> prev;
    size_t have = 0;
    bool partial = false;
    if (read_lines(op.c_str(), prev))
        for (const std::string &l : prev) {
            if (!l.empty() && l.back() == '\n') have++;
            else partial = true;
        }
    if (have > lines.size()) return PG_EINVAL;
    if (skipped) *skipped = (uint32_t)have;
    if (partial) {
        FILE *w = fopen(op.c_str(), "wb");
        if (!w) return PG_EINVAL;
        for (size_t i = 0; i < have; i++) fputs(prev[i].c_str(), w);
        if (fclose(w) != 0) return PG_EINVAL;
    }
    FILE *f = fopen(op.c_str(), have ? "ab" : "wb");
    if (!f) return PG_EINVAL;
    uint32_t nDone = 0;
    int rc = PG_OK;
    std::vector<pg_pos48> pos;
    std::vector<uint64_t> res;
    for (size_t first = have; first < lines.size() && rc == PG_OK; first += chunk) {
        const size_t n = std::min<size_t>(chunk, lines.size() - first);
        pos.resize(n);
        res.resize(n);
        for (size_t i = 0; i < n && rc == PG_OK; i++) rc = pg_fen_to_pos(lines[first + i].c_str(), &pos[i]);
        if (rc != PG_OK) break;
        rc = pg_perft_batch(ctx, pos.data(), (uint32_t)n, depth, slider, 1, 0, tt, res.data(), nullptr);
        if (rc != PG_OK) break;
        for (size_t i = 0; i < n; i++) {
            std::string line = lines[first + i];
            strip_newline(line);
            const unsigned long long occ = occurrence(line);
            fprintf(f, "%s %llu %llu\n", line.c_str(), (unsigned long long)res[i], (unsigned long long)res[i] * occ);
        }
        if (fflush(f) != 0) rc = PG_EINVAL;
        nDone += (uint32_t)n;
    }
    if (fclose(f) != 0 && rc == PG_OK) rc = PG_EINVAL;
    if (done) *done = nDone;
    return rc;
}

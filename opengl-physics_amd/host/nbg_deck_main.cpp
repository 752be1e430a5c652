// nbg_deck -- command-line deck runner on the GPU backend; the `--deck <file> --out <dir>
// [--frames N]` mode of projects/02_nbody_gravity/src/main.cpp:397-527.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ngrav_host.h"

int main(int argc, char** argv) {
    std::string deck, out = "out";
    int frames = 0, quiet = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--deck" && i + 1 < argc) deck = argv[++i];
        else if (a == "--out" && i + 1 < argc) out = argv[++i];
        else if (a == "--frames" && i + 1 < argc) frames = std::atoi(argv[++i]);
        else if (a == "--quiet") quiet = 1;
        else {
            std::fprintf(stderr, "usage: nbg_deck --deck <file.toml> [--out <dir>] [--frames N] [--quiet]\n");
            return 2;
        }
    }
    if (deck.empty()) {
        std::fprintf(stderr, "nbg_deck: --deck is required\n");
        return 2;
    }
    char err[512] = {0};
    if (ngh_run_deck(deck.c_str(), out.c_str(), frames, quiet, err, sizeof(err)) != 0) {
        std::fprintf(stderr, "nbg_deck: %s\n", err);
        return 1;
    }
    return 0;
}

// cli_terawarp.cpp -- command line with the reference's -i / -o / -p layout:
//   terawarp -i <image> <marker_sub> <marker_tar> -o <out_dir> -p <params...>
// (the Vaa3D invocation `v3d -x <plugin> -f bigimagewarp -i ... -o ... -p ...`
// of plugin_Bigimagewarp.cpp:109-120 without Vaa3D).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/terawarp_host.h"

int main(int argc, char** argv) {
    std::vector<const char*> in, out, par;
    std::vector<const char*>* cur = nullptr;
    std::string func = "bigimagewarp";
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "-i")) cur = &in;
        else if (!std::strcmp(argv[i], "-o")) cur = &out;
        else if (!std::strcmp(argv[i], "-p")) cur = &par;
        else if (!std::strcmp(argv[i], "-f") && i + 1 < argc) func = argv[++i], cur = nullptr;
        else if (!std::strcmp(argv[i], "-h") || !std::strcmp(argv[i], "--help")) func = "help";
        else if (cur) cur->push_back(argv[i]);
        else {
            std::fprintf(stderr, "unexpected argument %s\n", argv[i]);
            return 2;
        }
    }
    const int ok = tw_plugin_dofunc(func.c_str(), in.data(), (int)in.size(), par.data(), (int)par.size(),
                                    out.data(), (int)out.size());
    if (!ok) {
        std::fprintf(stderr, "terawarp: %s\n", tw_io_last_error());
        return 1;
    }
    return 0;
}

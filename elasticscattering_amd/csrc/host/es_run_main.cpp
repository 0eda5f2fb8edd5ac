// es_run — command line of the host runner: `es_run <config>` runs the
// simulation a .config describes (src/main.cpp:31-93 without the Debug-only
// test mode; `es_run help` prints the usage).
#include <cstdio>
#include <cstring>

#include "../../../include/es_runner.h"

int main(int argc, char** argv)
{
    const char* cfg = "default.config";
    if (argc > 1) {
        if (std::strcmp(argv[1], "help") == 0 || std::strcmp(argv[1], "--help") == 0) {
            std::printf("usage: es_run [config file]   (default: default.config)\n");
            return 0;
        }
        cfg = argv[1];
    }
    const int rc = es_run(cfg, 1, nullptr, 0);
    return rc == 0 ? 0 : 1;
}

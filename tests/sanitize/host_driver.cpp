// Test driver for the sanitizer builds of the host runner (es_config.cpp,
// es_runner.cpp, es_png.cpp): parses every config named on the command line
// through es_config_parse, writes one PNG through es_log_images, and runs
// es_run on the first config, which on a machine without a GPU ends in the
// solver's no-device error.  Prints what it got, one line per call.
//   host_driver [--no-run] OUT_DIR CONFIG...   (--no-run: skip es_run)
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/es_abi.h"
#include "../../include/es_runner.h"

int main(int argc, char** argv)
{
    int a = 1;
    const bool run = !(argc > 1 && std::string(argv[1]) == "--no-run");
    if (!run) a++;
    if (argc < a + 2) return 2;
    const std::string out = argv[a++];
    std::vector<char> buf(1 << 16);
    for (int i = a; i < argc; i++) {
        const int rc = es_config_parse(argv[i], buf.data(), (int)buf.size());
        std::printf("parse %s rc=%d %s\n", argv[i], rc, rc >= -1 ? buf.data() : "(buffer too small)");
        // a buffer too small for the JSON: the size it needs, nothing written
        char small[4] = {'x', 'x', 'x', 0};
        const int need = es_config_parse(argv[i], small, 4);
        std::printf("small rc=%d untouched=%d\n", need < 0 ? -1 : need, small[0] == 'x');
    }
    const int dim = 7;
    std::vector<double> lt(dim * dim), xx(dim * dim), xy(dim * dim);
    for (int k = 0; k < dim * dim; k++) {
        lt[k] = (k % 5 == 0) ? 0.0 : 1e-12 * (1 + k);
        xx[k] = std::sin(0.3 * k) * 1e-3;
        xy[k] = (k % 7 == 0) ? 0.0 : std::cos(0.2 * k) * 1e-4;
    }
    std::printf("png rc=%d\n", es_log_images((out + "/img.png").c_str(), dim, lt.data(), xx.data(), xy.data()));
    if (run) {
        char dir[512] = {0};
        std::printf("run rc=%d\n", es_run(argv[a], 0, dir, (int)sizeof dir));
    }
    return 0;
}

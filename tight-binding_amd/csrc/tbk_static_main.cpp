// tbk_static_main.cpp -- the native program for this path: tbk_static DIR
// reads DIR/SKIN (and the parameter, k-point and space-group files it
// names), solves every structure on the GPU and writes DIR/SKENG, as the
// reference's program static does (static.f) with bande on the device.
// Options: --mixed (mixed precision), --device N.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "tbk.h"

int main(int argc, char **argv)
{
    const char *dir = ".";
    int device = 0, precision = TBK_FP64;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--mixed")) precision = TBK_MIXED;
        else if (!std::strcmp(argv[i], "--device") && i + 1 < argc) device = std::atoi(argv[++i]);
        else dir = argv[i];
    }
    int n = 0;
    const int rc = tbk_static_run(dir, device, precision, &n);
    std::printf("tbk_static: %d structures, %s\n", n, tbk_error_string(rc));
    return rc == TBK_OK ? 0 : 1;
}

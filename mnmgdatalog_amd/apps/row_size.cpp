// row_size -- rows of an int32 pair file, as the reference's row_size.cpp:
// prints "<file>: <size / 8>" (0 when the file cannot be read).
#include <cstdio>
#include <sys/stat.h>

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: row_size <file.bin>\n");
        return 1;
    }
    struct stat st {};
    long long size = stat(argv[1], &st) == 0 ? (long long)st.st_size : 0;
    std::printf("%s: %lld\n", argv[1], size / (long long)(2 * sizeof(int)));
    return 0;
}

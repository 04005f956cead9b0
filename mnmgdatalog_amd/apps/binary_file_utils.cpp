// binary_file_utils -- the reference's format converter (binary_file_utils.py:1-54)
// as a native tool:
//   binary_file_utils txt_to_bin <in.txt> <out.bin>
//     each line: tabs read as spaces, whitespace-separated fields, the fields
//     that are integers kept, the first two written as native int32 (a line
//     with one integer writes one, with none writes nothing)
//   binary_file_utils bin_to_txt <in.bin> <out.txt>
//     8-byte rows written as "<a>\t<b>\n"
// Anything else prints the usage.  A value outside int32, an unreadable
// input or a trailing partial row in a .bin input is an error (the
// reference's script stops with an exception there), with a non-zero exit.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// A field as the script's int() takes it: optional sign, decimal digits with
// single underscores between them.
static bool parse_int(const std::string &f, long long *out, bool *range_ok) {
    size_t i = 0;
    bool neg = false;
    if (i < f.size() && (f[i] == '+' || f[i] == '-')) neg = f[i++] == '-';
    if (i >= f.size()) return false;
    unsigned long long v = 0;
    bool digit_last = false, any = false, over = false;
    for (; i < f.size(); i++) {
        char c = f[i];
        if (c >= '0' && c <= '9') {
            v = v * 10 + (unsigned)(c - '0');
            if (v > (1ull << 40)) over = true;
            digit_last = any = true;
        } else if (c == '_' && digit_last && i + 1 < f.size()) {
            digit_last = false;
        } else {
            return false;
        }
    }
    if (!any || !digit_last) return false;
    long long s = neg ? -(long long)v : (long long)v;
    *range_ok = !over && s >= INT32_MIN && s <= INT32_MAX;
    *out = s;
    return true;
}

static int txt_to_bin(const char *in, const char *out) {
    FILE *fi = std::fopen(in, "r");
    if (!fi) {
        std::fprintf(stderr, "cannot open %s: %s\n", in, std::strerror(errno));
        return 1;
    }
    FILE *fo = std::fopen(out, "wb");
    if (!fo) {
        std::fprintf(stderr, "cannot open %s: %s\n", out, std::strerror(errno));
        std::fclose(fi);
        return 1;
    }
    std::string line;
    int ch;
    bool eof = false;
    while (!eof) {
        line.clear();
        while ((ch = std::fgetc(fi)) != EOF && ch != '\n') line.push_back((char)ch);
        eof = ch == EOF;
        if (eof && line.empty()) break;
        std::vector<int32_t> row;
        size_t i = 0;
        while (i < line.size() && row.size() < 2) {
            while (i < line.size() && std::strchr(" \t\r\v\f", line[i])) i++;
            size_t j = i;
            while (j < line.size() && !std::strchr(" \t\r\v\f", line[j])) j++;
            if (j > i) {
                long long v;
                bool ok;
                if (parse_int(line.substr(i, j - i), &v, &ok)) {
                    if (!ok) {
                        std::fprintf(stderr, "value %s does not fit int32\n", line.substr(i, j - i).c_str());
                        std::fclose(fi);
                        std::fclose(fo);
                        return 1;
                    }
                    row.push_back((int32_t)v);
                }
            }
            i = j;
        }
        if (!row.empty() && std::fwrite(row.data(), sizeof(int32_t), row.size(), fo) != row.size()) {
            std::fprintf(stderr, "write error on %s\n", out);
            std::fclose(fi);
            std::fclose(fo);
            return 1;
        }
    }
    std::fclose(fi);
    if (std::fclose(fo) != 0) return 1;
    std::printf("Converted %s to binary file: %s\n", in, out);
    return 0;
}

static int bin_to_txt(const char *in, const char *out) {
    FILE *fi = std::fopen(in, "rb");
    if (!fi) {
        std::fprintf(stderr, "cannot open %s: %s\n", in, std::strerror(errno));
        return 1;
    }
    FILE *fo = std::fopen(out, "w");
    if (!fo) {
        std::fprintf(stderr, "cannot open %s: %s\n", out, std::strerror(errno));
        std::fclose(fi);
        return 1;
    }
    int32_t r[2];
    size_t got;
    int rc = 0;
    while ((got = std::fread(r, 1, sizeof r, fi)) == sizeof r) std::fprintf(fo, "%d\t%d\n", r[0], r[1]);
    if (got != 0) {
        std::fprintf(stderr, "%s: trailing %zu bytes are not a whole row\n", in, got);
        rc = 1;
    }
    std::fclose(fi);
    if (std::fclose(fo) != 0) rc = 1;
    if (rc == 0) std::printf("Converted %s to txt file: %s\n", in, out);
    return rc;
}

int main(int argc, char **argv) {
    if (argc != 4) {
        std::printf("Usage: binary_file_utils <conversion_type> <input_file_path> <output_file_path>\n");
        std::printf("Supported conversion types: txt_to_bin, bin_to_txt\n");
        return 0;
    }
    if (!std::strcmp(argv[1], "txt_to_bin")) return txt_to_bin(argv[2], argv[3]);
    if (!std::strcmp(argv[1], "bin_to_txt")) return bin_to_txt(argv[2], argv[3]);
    std::printf("Invalid conversion type. Supported types: txt_to_bin, bin_to_txt\n");
    return 0;
}

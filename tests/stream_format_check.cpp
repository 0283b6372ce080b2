// Writes numbers through C++ streams set up the way GLAC's writers set them
// (default format for header values, std::fixed + setprecision(15) for
// rows), one file per mode, so the Python writers can be compared with the
// C++ library's formatting byte for byte.
//   usage: stream_format_check <out_dir> v0 v1 ...
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <string>

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    std::ofstream g(dir + "/default.txt"), f(dir + "/fixed15.txt"), s(dir + "/to_string.txt");
    f << std::fixed << std::setprecision(15);
    for (int i = 2; i < argc; i++) {
        const double v = std::strtod(argv[i], nullptr);
        g << v << "\n";
        f << v << "\n";
        s << std::to_string(v) << "\n";
    }
    return 0;
}

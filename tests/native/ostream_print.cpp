// Prints values through a default std::ostream in the reference's text
// forms: "[v0, v1, ...]'" for a vector and "index (x, y, z)" for a point.
// Input on stdin: a count, then that many doubles (as hex floats), then a
// point index and three more doubles.
#include <cstdio>
#include <iostream>
#include <vector>

int main()
{
	int n = 0;
	if (std::scanf("%d", &n) != 1) return 1;
	std::vector<double> v(n);
	for (int i = 0; i < n; ++i)
		if (std::scanf("%la", &v[i]) != 1) return 1;
	unsigned idx = 0;
	double p[3];
	if (std::scanf("%u %la %la %la", &idx, &p[0], &p[1], &p[2]) != 4) return 1;
	std::cout << "[";
	for (int i = 0; i < n; ++i) {
		if (i) std::cout << ", ";
		std::cout << v[i];
	}
	std::cout << "]'" << "\n";
	std::cout << idx << " (" << p[0] << ", " << p[1] << ", " << p[2] << ")" << "\n";
	return 0;
}

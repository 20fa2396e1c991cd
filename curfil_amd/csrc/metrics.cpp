// Prediction quality: calculatePixelAccuracy and ConfusionMatrix
// (predict.cpp:44-200), on host label images.
#include <vector>

#include "cf_host.h"

namespace {

bool ignoredLabel(int label, const int* ignored, int numIgnored) {
    for (int i = 0; i < numIgnored; i++) {
        if (ignored[i] == label) return true;
    }
    return false;
}

}  // namespace

extern "C" {

int cf_pixel_accuracy(const uint8_t* prediction, const uint8_t* groundTruth, size_t n, int includeVoid,
        const int* ignored, int numIgnored, double* confusion, int numClasses, double* accuracy) {
    if (!prediction || !groundTruth || !accuracy || (numIgnored > 0 && !ignored)) return CF_ERR_ARG;
    if (confusion && numClasses <= 0) return CF_ERR_ARG;
    size_t correct = 0, wrong = 0;
    for (size_t i = 0; i < n; i++) {
        const int label = groundTruth[i];
        // only the accuracy without void skips the ignored labels (predict.cpp:170-180)
        if (!includeVoid && ignoredLabel(label, ignored, numIgnored)) continue;
        const int predicted = prediction[i];
        if (predicted == label) {
            correct++;
        } else {
            wrong++;
        }
        if (confusion) {
            if (label >= numClasses || predicted >= numClasses) return CF_ERR_ARG;
            confusion[static_cast<size_t>(label) * numClasses + predicted] += 1.0;
        }
    }
    const size_t pixels = includeVoid ? n : correct + wrong;
    *accuracy = static_cast<double>(correct) / static_cast<double>(pixels);
    return CF_OK;
}

int cf_confusion_normalize(double* confusion, int numClasses) {
    if (!confusion || numClasses <= 0) return CF_ERR_ARG;
    // ConfusionMatrix::normalize (predict.cpp:102-140): rows that sum to zero stay zero
    for (int l = 0; l < numClasses; l++) {
        double* row = confusion + static_cast<size_t>(l) * numClasses;
        double sum = 0.0;
        for (int p = 0; p < numClasses; p++) sum += row[p];
        if (sum == 0.0) continue;
        for (int p = 0; p < numClasses; p++) row[p] /= sum;
    }
    return CF_OK;
}

int cf_confusion_class_accuracy(const double* confusion, int numClasses, int normalized, int includeVoid,
        const int* ignored, int numIgnored, double* out) {
    if (!confusion || numClasses <= 0 || !out || (numIgnored > 0 && !ignored)) return CF_ERR_ARG;
    // ConfusionMatrix::averageClassAccuracy (predict.cpp:71-100) of the row-normalised matrix
    std::vector<double> m(confusion, confusion + static_cast<size_t>(numClasses) * numClasses);
    if (!normalized) cf_confusion_normalize(m.data(), numClasses);
    double sum = 0.0;
    size_t count = 0;
    for (int l = 0; l < numClasses; l++) {
        if (!includeVoid && ignoredLabel(l, ignored, numIgnored)) continue;
        sum += m[static_cast<size_t>(l) * numClasses + l];
        count++;
    }
    *out = count ? sum / static_cast<double>(count) : 0.0;
    return CF_OK;
}

}  // extern "C"
